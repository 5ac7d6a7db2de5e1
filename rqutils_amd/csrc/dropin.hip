// Operator-level drop-ins of rqutils/sqd.py that take the reference's own
// arrays: packed padded state rows (uniquify_states, sqd.py:414-445), cached
// source indices, diagonals or sign bits (apply_h and its five cached
// siblings, sqd.py:590-684; get_diag_signs / compute_diagonal, sqd.py:
// 506-551).
//
// A packed row of b <= 8 bytes is read as one big-endian integer, its "row
// key": byte order is the reference's lexicographic row order (fill rows of
// 0xFF sort last), XOR with a packed X signature is the reference's S ^ x,
// and AND with a packed Z signature has the reference's popcount. So every
// kernel here works on row keys as they are, pad bit included.
//
// These functions exist for callers that hold the reference's arrays; the
// fast product is sqdk_matvec (bucket join) or sqdk_cache_matvec (hit
// table).
#include "internal.hpp"

namespace sqdk {

namespace {

__global__ void __launch_bounds__(256)
rows_to_keys_kernel(const uint8_t* __restrict__ rows, int64_t n, int bp, uint64_t* __restrict__ keys) {
  const int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x;
  if (i >= n) return;
  const uint8_t* r = rows + i * bp;
  uint64_t k = 0;
  for (int b = 0; b < bp; ++b) k = (k << 8) | r[b];
  keys[i] = k;
}

__global__ void __launch_bounds__(256)
keys_to_rows_kernel(const uint64_t* __restrict__ keys, int64_t n_keys, int64_t n_rows, int bp,
                    uint8_t* __restrict__ rows) {
  const int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x;
  if (i >= n_rows) return;
  uint8_t* r = rows + i * bp;
  if (i < n_keys) {
    const uint64_t k = keys[i];
    for (int b = 0; b < bp; ++b) r[b] = (uint8_t)(k >> (8 * (bp - 1 - b)));
  } else {
    for (int b = 0; b < bp; ++b) r[b] = 0xFF;
  }
}

// Position of t among keys[0, n) (ascending), or -1.
__device__ inline int64_t find_key(const uint64_t* __restrict__ keys, int64_t n, uint64_t t) {
  int64_t lo = 0, len = n;
  while (len > 0) {
    const int64_t half = len >> 1;
    if (keys[lo + half] < t) {
      lo += half + 1;
      len -= half + 1;
    } else {
      len = half;
    }
  }
  return (lo < n && keys[lo] == t) ? lo : -1;
}

__device__ inline double2 load_v(const double2* v, int64_t s) { return v[s]; }
__device__ inline double2 load_v(const float2* v, int64_t s) {
  const float2 a = v[s];
  return make_double2(a.x, a.y);
}
__device__ inline void store_v(double2* o, int64_t i, double2 a) { o[i] = a; }
__device__ inline void store_v(float2* o, int64_t i, double2 a) {
  o[i] = make_float2((float)a.x, (float)a.y);
}

// out[i] = sum_g C_g(s_i) v[source_g(i)] in group order, FP64 sums (the
// reference's lax.scan over groups, sqd.py:604-605, 636-637, 653-654).
// Rows at or past n_real are fill rows (outside the subspace): out = 0.
template <typename VT, bool CACHED, int FORM>
__global__ void __launch_bounds__(256)
groups_matvec_kernel(sqdk_groups g, const VT* __restrict__ v, VT* __restrict__ out) {
  const int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x;
  if (i >= g.n) return;
  if (i >= g.n_real) {
    store_v(out, i, make_double2(0.0, 0.0));
    return;
  }
  const uint64_t key = g.keys[i];
  const double2* coef = reinterpret_cast<const double2*>(g.coef);
  double2 acc = make_double2(0.0, 0.0);
  for (int j = 0; j < g.n_groups; ++j) {
    int64_t s;
    if (CACHED)
      s = g.xsource[(int64_t)j * g.ld + i];
    else
      s = find_key(g.keys, g.n_real, key ^ g.xmask[j]);
    if (s < 0 || s >= g.n_real) continue;
    double2 c = make_double2(0.0, 0.0);
    if (FORM == SQDK_COEF_TERMS) {
      for (int k = 0; k < g.kmax; ++k) {
        const double2 ck = coef[(int64_t)j * g.kmax + k];
        const bool odd = (__popcll(g.zmask[(int64_t)j * g.kmax + k] & key) & 1) != 0;
        c.x += odd ? -ck.x : ck.x;
        c.y += odd ? -ck.y : ck.y;
      }
    } else if (FORM == SQDK_COEF_DIAG) {
      c = reinterpret_cast<const double2*>(g.diag)[(int64_t)j * g.ld + i];
    } else {
      const uint8_t* sb = g.signs + ((int64_t)j * g.ld + i) * g.sign_bytes;
      for (int k = 0; k < g.kmax; ++k) {
        const double2 ck = coef[(int64_t)j * g.kmax + k];
        const bool odd = ((sb[k >> 3] >> (7 - (k & 7))) & 1) != 0;
        c.x += odd ? -ck.x : ck.x;
        c.y += odd ? -ck.y : ck.y;
      }
    }
    acc = cadd(acc, cmul(c, load_v(v, s)));
  }
  store_v(out, i, acc);
}

// Sign bits of one group's terms: bit 7 - k % 8 of byte k / 8 of row i is
// the parity of z_k & s_i (get_diag_signs, sqd.py:506-528).
__global__ void __launch_bounds__(256)
diag_signs_kernel(const uint64_t* __restrict__ zmask, int nt, const uint64_t* __restrict__ keys,
                  int64_t n, int nb, uint8_t* __restrict__ out) {
  const int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x;
  if (i >= n) return;
  const uint64_t key = keys[i];
  for (int b = 0; b < nb; ++b) {
    uint32_t byte = 0;
    for (int k = 8 * b; k < min(nt, 8 * b + 8); ++k)
      byte |= (uint32_t)(__popcll(zmask[k] & key) & 1) << (7 - (k & 7));
    out[i * nb + b] = (uint8_t)byte;
  }
}

// C[i] = sum_k c_k (1 - 2 bit_k(i)) with bit k read at position 7 - k % 8 of
// byte k / 8 (compute_diagonal, sqd.py:531-551, which reads bit 7 - (k & 255)
// and so gives +1 for every term from the ninth on).
__global__ void __launch_bounds__(256)
compute_diagonal_kernel(const uint8_t* __restrict__ signs, int64_t n, int nb,
                        const double2* __restrict__ coef, int nt, double2* __restrict__ out) {
  const int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x;
  if (i >= n) return;
  double2 c = make_double2(0.0, 0.0);
  for (int k = 0; k < nt; ++k) {
    const bool odd = ((signs[i * nb + (k >> 3)] >> (7 - (k & 7))) & 1) != 0;
    c.x += odd ? -coef[k].x : coef[k].x;
    c.y += odd ? -coef[k].y : coef[k].y;
  }
  out[i] = c;
}

template <typename VT>
int launch_groups(const sqdk_groups* g, const void* v, void* out, hipStream_t s) {
  const unsigned blocks = (unsigned)ceil_div(g->n, 256);
  const VT* vv = static_cast<const VT*>(v);
  VT* oo = static_cast<VT*>(out);
  const bool cached = g->xsource != nullptr;
#define SQDK_GROUPS_CASE(C, F)                                                    \
  if (cached == C && g->coef_form == F) {                                        \
    groups_matvec_kernel<VT, C, F><<<blocks, 256, 0, s>>>(*g, vv, oo);           \
    SQDK_LAUNCHED("groups_matvec_kernel");                                       \
    return SQDK_OK;                                                              \
  }
  SQDK_GROUPS_CASE(true, SQDK_COEF_TERMS)
  SQDK_GROUPS_CASE(true, SQDK_COEF_DIAG)
  SQDK_GROUPS_CASE(true, SQDK_COEF_SIGNS)
  SQDK_GROUPS_CASE(false, SQDK_COEF_TERMS)
  SQDK_GROUPS_CASE(false, SQDK_COEF_DIAG)
  SQDK_GROUPS_CASE(false, SQDK_COEF_SIGNS)
#undef SQDK_GROUPS_CASE
  return fail(SQDK_EINVAL, "unknown coefficient form %d", g->coef_form);
}

}  // namespace

}  // namespace sqdk

using namespace sqdk;

extern "C" {

int sqdk_rows_to_keys(const uint8_t* rows, int64_t n, int32_t row_bytes, uint64_t* keys,
                      void* stream) {
  SQDK_REQUIRE(n >= 0, "negative row count");
  SQDK_REQUIRE(row_bytes >= 1 && row_bytes <= 8, "rows of %d bytes (1..8 supported)", row_bytes);
  if (n == 0) return SQDK_OK;
  SQDK_REQUIRE(rows && keys, "NULL pointer argument");
  rows_to_keys_kernel<<<(unsigned)ceil_div(n, 256), 256, 0, as_stream(stream)>>>(rows, n,
                                                                                 row_bytes, keys);
  SQDK_LAUNCHED("rows_to_keys_kernel");
  return SQDK_OK;
}

int sqdk_keys_to_rows(const uint64_t* keys, int64_t n_keys, int64_t n_rows, int32_t row_bytes,
                      uint8_t* rows, void* stream) {
  SQDK_REQUIRE(n_keys >= 0 && n_rows >= 0, "negative count");
  SQDK_REQUIRE(row_bytes >= 1 && row_bytes <= 8, "rows of %d bytes (1..8 supported)", row_bytes);
  if (n_rows == 0) return SQDK_OK;
  SQDK_REQUIRE(rows && (keys || n_keys == 0), "NULL pointer argument");
  keys_to_rows_kernel<<<(unsigned)ceil_div(n_rows, 256), 256, 0, as_stream(stream)>>>(
      keys, n_keys, n_rows, row_bytes, rows);
  SQDK_LAUNCHED("keys_to_rows_kernel");
  return SQDK_OK;
}

int sqdk_groups_matvec(const sqdk_groups* g, int32_t precision, const void* v, void* out,
                       void* stream) {
  SQDK_REQUIRE(g, "NULL pointer argument");
  SQDK_REQUIRE(g->n >= 0 && g->n_real >= 0 && g->n_real <= g->n, "bad row counts");
  SQDK_REQUIRE(g->n < ((int64_t)1 << 31), "at most 2^31 - 1 rows");
  SQDK_REQUIRE(g->n_groups >= 0, "negative group count");
  SQDK_REQUIRE(precision == SQDK_C128 || precision == SQDK_C64, "unknown precision %d",
               precision);
  if (g->n == 0) return SQDK_OK;
  SQDK_REQUIRE(g->keys && v && out, "NULL pointer argument");
  if (g->n_groups > 0) {
    SQDK_REQUIRE(g->xsource || g->xmask, "neither source indices nor X signatures");
    SQDK_REQUIRE(!g->xsource || g->ld >= g->n, "source stride %lld below %lld rows",
                 (long long)g->ld, (long long)g->n);
    if (g->coef_form == SQDK_COEF_TERMS)
      SQDK_REQUIRE(g->kmax == 0 || (g->zmask && g->coef), "terms need zmask and coef");
    else if (g->coef_form == SQDK_COEF_DIAG)
      SQDK_REQUIRE(g->diag && g->ld >= g->n, "diagonals need diag and ld >= n");
    else if (g->coef_form == SQDK_COEF_SIGNS)
      SQDK_REQUIRE(g->signs && g->ld >= g->n && (g->kmax == 0 || g->coef) &&
                       g->sign_bytes * 8 >= g->kmax,
                   "sign bits need signs, coef, ld >= n and 8 sign_bytes >= kmax");
  }
  hipStream_t s = as_stream(stream);
  return precision == SQDK_C128 ? launch_groups<double2>(g, v, out, s)
                                : launch_groups<float2>(g, v, out, s);
}

int sqdk_diag_signs(const uint64_t* zmask, int32_t n_terms, const uint64_t* keys, int64_t n,
                    int32_t sign_bytes, uint8_t* out, void* stream) {
  SQDK_REQUIRE(n >= 0 && n_terms >= 0, "negative count");
  SQDK_REQUIRE(sign_bytes * 8 >= n_terms && sign_bytes >= 1, "sign_bytes %d for %d terms",
               sign_bytes, n_terms);
  if (n == 0) return SQDK_OK;
  SQDK_REQUIRE(keys && out && (zmask || n_terms == 0), "NULL pointer argument");
  diag_signs_kernel<<<(unsigned)ceil_div(n, 256), 256, 0, as_stream(stream)>>>(
      zmask, n_terms, keys, n, sign_bytes, out);
  SQDK_LAUNCHED("diag_signs_kernel");
  return SQDK_OK;
}

int sqdk_compute_diagonal(const uint8_t* signs, int64_t n, int32_t sign_bytes, const double* coef,
                          int32_t n_terms, double* out, void* stream) {
  SQDK_REQUIRE(n >= 0 && n_terms >= 0, "negative count");
  SQDK_REQUIRE(sign_bytes * 8 >= n_terms, "sign_bytes %d for %d terms", sign_bytes, n_terms);
  if (n == 0) return SQDK_OK;
  SQDK_REQUIRE(signs && out && (coef || n_terms == 0), "NULL pointer argument");
  compute_diagonal_kernel<<<(unsigned)ceil_div(n, 256), 256, 0, as_stream(stream)>>>(
      signs, n, sign_bytes, reinterpret_cast<const double2*>(coef), n_terms,
      reinterpret_cast<double2*>(out));
  SQDK_LAUNCHED("compute_diagonal_kernel");
  return SQDK_OK;
}

}  // extern "C"
