// Subspace preparation and the operator-level lookups of the SQD path:
// key validation, bucket / fine directories, row chunks, the bit-exact
// source-index search (rqutils/sqd.py:448-503) and per-group coefficient
// vectors (rqutils/sqd.py:554-575).
#include "common.hpp"

namespace sqdk {

std::string& last_error_slot() {
  static thread_local std::string slot;
  return slot;
}

namespace {

constexpr int kBlock = 256;
// Directory entries a fine directory may hold (its kernels loop over them
// with a grid of at most 2^20 workgroups).
constexpr int64_t kMaxFineEntries = 1ll << 31;

inline unsigned grid_stride_blocks(int64_t total) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(total, kBlock), 1ll << 20));
}

// Strict ascent, key width and popcount presence. One atomicOr per block.
__global__ void __launch_bounds__(kBlock)
scan_keys_kernel(const uint64_t* __restrict__ keys, int64_t n, int nbits,
                 uint64_t* __restrict__ pc_mask, int32_t* __restrict__ err) {
  __shared__ unsigned long long s_pc[2];
  __shared__ int s_err;
  if (threadIdx.x < 2) s_pc[threadIdx.x] = 0;
  if (threadIdx.x == 0) s_err = 0;
  __syncthreads();
  unsigned long long pc0 = 0, pc1 = 0;
  int bad = 0;
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock) {
    const uint64_t k = keys[i];
    if (nbits < 64 && (k >> nbits) != 0) ++bad;
    if (i > 0 && keys[i - 1] >= k) ++bad;
    const int w = __popcll(k);
    if (w < 64) pc0 |= 1ull << w; else pc1 |= 1ull;
  }
  if (pc0) atomicOr(&s_pc[0], pc0);
  if (pc1) atomicOr(&s_pc[1], pc1);
  if (bad) atomicAdd(&s_err, bad);
  __syncthreads();
  if (threadIdx.x == 0) {
    if (s_pc[0]) atomicOr(reinterpret_cast<unsigned long long*>(pc_mask), s_pc[0]);
    if (s_pc[1]) atomicOr(reinterpret_cast<unsigned long long*>(pc_mask) + 1, s_pc[1]);
    if (s_err) atomicAdd(err, s_err);
  }
}

__global__ void zero_scan_outputs(uint64_t* pc_mask, int32_t* err) {
  if (threadIdx.x == 0) {
    pc_mask[0] = 0;
    pc_mask[1] = 0;
    err[0] = 0;
  }
}

// dir[p] = lower_bound(keys, p << shift), p in [0, np]; the last entry is n.
__global__ void __launch_bounds__(kBlock)
dir_kernel(const uint64_t* __restrict__ keys, int64_t n, int shift, int64_t np,
           uint32_t* __restrict__ dir) {
  const int64_t p = blockIdx.x * (int64_t)kBlock + threadIdx.x;
  if (p > np) return;
  if (p == np) {
    dir[p] = (uint32_t)n;
    return;
  }
  const uint64_t t = (uint64_t)p << shift;
  dir[p] = (uint32_t)lower_bound_range<uint64_t>(keys, 0, n, t);
}

// fdir[(b << d) + e] = (first key of bucket b with sub-bucket >= e) - bstart[b].
__global__ void __launch_bounds__(kBlock)
fine_kernel(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ bstart,
            int m, int d, int64_t total, uint16_t* __restrict__ fdir) {
  for (int64_t idx = blockIdx.x * (int64_t)kBlock + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * kBlock) {
    const int64_t b = idx >> d;
    const uint64_t e = (uint64_t)(idx & ((1ll << d) - 1));
    const int64_t lo = bstart[b];
    const int64_t hi = bstart[b + 1];
    const uint64_t t = ((uint64_t)b << m) | (e << (m - d));
    const int64_t j = lower_bound_range<uint64_t>(keys, lo, hi, t);
    fdir[idx] = (uint16_t)(j - lo);
  }
}

// fbits[e] = OR of 1 << (key & 31) over the keys of 32-value sub-bucket e
// (fine directory with d = m - 5); one thread per sub-bucket, no atomics.
__global__ void __launch_bounds__(kBlock)
bits_kernel(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ bstart,
            const uint16_t* __restrict__ fdir, int d, int64_t total, uint32_t* __restrict__ fbits) {
  for (int64_t idx = blockIdx.x * (int64_t)kBlock + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * kBlock) {
    const int64_t b = idx >> d;
    const int64_t e = idx & ((1ll << d) - 1);
    const int64_t base = bstart[b];
    const int64_t lo = base + fdir[idx];
    const int64_t hi = (e + 1 < (1ll << d)) ? base + fdir[idx + 1] : (int64_t)bstart[b + 1];
    uint32_t bits = 0;
    for (int64_t j = lo; j < hi; ++j) bits |= 1u << (uint32_t)(keys[j] & 31u);
    fbits[idx] = bits;
  }
}

// Row chunks. Single block: serial scan over buckets in block-sized strides.
__global__ void __launch_bounds__(kBlock)
items_kernel(const uint32_t* __restrict__ bstart, int64_t nb, int chunk,
             uint32_t* __restrict__ items, int64_t max_items,
             int32_t* __restrict__ n_items) {
  __shared__ int64_t s_scan[kBlock];
  __shared__ int64_t s_base;
  if (threadIdx.x == 0) s_base = 0;
  __syncthreads();
  for (int64_t b0 = 0; b0 < nb; b0 += kBlock) {
    const int64_t b = b0 + threadIdx.x;
    int64_t cnt = 0;
    uint32_t lo = 0, hi = 0;
    if (b < nb) {
      lo = bstart[b];
      hi = bstart[b + 1];
      cnt = ((int64_t)hi - lo + chunk - 1) / chunk;
    }
    s_scan[threadIdx.x] = cnt;
    __syncthreads();
    // Hillis-Steele inclusive scan over the block.
    for (int off = 1; off < kBlock; off <<= 1) {
      int64_t add = threadIdx.x >= off ? s_scan[threadIdx.x - off] : 0;
      __syncthreads();
      s_scan[threadIdx.x] += add;
      __syncthreads();
    }
    const int64_t first = s_base + s_scan[threadIdx.x] - cnt;
    for (int64_t k = 0; k < cnt; ++k) {
      if (first + k < max_items) items[first + k] = lo + (uint32_t)(k * chunk);
    }
    __syncthreads();
    if (threadIdx.x == kBlock - 1) s_base += s_scan[kBlock - 1];
    __syncthreads();
  }
  if (threadIdx.x == 0) n_items[0] = (int32_t)s_base;
}

__global__ void __launch_bounds__(kBlock)
xsource_kernel(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ bstart,
               int64_t n, int nbits, int shift, uint64_t x, int32_t* __restrict__ out) {
  const int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x;
  if (i >= n) return;
  out[i] = find_key_global(keys, bstart, nbits, shift, keys[i] ^ x);
}

__global__ void __launch_bounds__(kBlock)
diagonal_kernel(const uint64_t* __restrict__ zmask, const double2* __restrict__ coef,
                int t0, int t1, const uint64_t* __restrict__ keys, int64_t n,
                double2* __restrict__ out) {
  const int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x;
  if (i >= n) return;
  double cr, ci;
  group_coef<double2>(zmask, coef, t0, t1, keys[i], cr, ci);
  out[i] = make_double2(cr, ci);
}

int check_space(const sqdk_space* sp) {
  SQDK_REQUIRE(sp != nullptr, "space is NULL");
  SQDK_REQUIRE(sp->keys && sp->bstart, "space has NULL keys or bucket directory");
  SQDK_REQUIRE(sp->n >= 1 && sp->n < (1ll << 31), "space size %lld out of range",
               (long long)sp->n);
  SQDK_REQUIRE(sp->num_qubits >= 1 && sp->num_qubits <= 64, "num_qubits %d out of range",
               sp->num_qubits);
  SQDK_REQUIRE(sp->bucket_bits >= 0 && sp->bucket_bits <= sp->num_qubits &&
                   sp->num_qubits - sp->bucket_bits <= 26,
               "bucket_bits %d invalid for %d qubits", sp->bucket_bits, sp->num_qubits);
  return SQDK_OK;
}

}  // namespace

int check_space_public(const sqdk_space* sp) { return check_space(sp); }

}  // namespace sqdk

using namespace sqdk;

extern "C" {

int sqdk_abi_version(void) { return SQDK_ABI_VERSION; }

const char* sqdk_last_error(void) { return last_error_slot().c_str(); }

int sqdk_space_scan(const uint64_t* keys, int64_t n, int32_t num_qubits,
                    uint64_t* pc_mask, int32_t* err_out, void* stream) {
  SQDK_REQUIRE(keys && pc_mask && err_out, "NULL pointer argument");
  SQDK_REQUIRE(n >= 1 && n < (1ll << 31), "n=%lld out of range", (long long)n);
  SQDK_REQUIRE(num_qubits >= 1 && num_qubits <= 64, "num_qubits %d out of range", num_qubits);
  hipStream_t s = as_stream(stream);
  zero_scan_outputs<<<1, 64, 0, s>>>(pc_mask, err_out);
  SQDK_LAUNCHED("zero_scan_outputs");
  const int64_t blocks = std::min<int64_t>(ceil_div(n, kBlock), 4096);
  scan_keys_kernel<<<(unsigned)blocks, kBlock, 0, s>>>(keys, n, num_qubits, pc_mask, err_out);
  SQDK_LAUNCHED("scan_keys_kernel");
  return SQDK_OK;
}

int sqdk_space_dir(const uint64_t* keys, int64_t n, int32_t num_qubits, int32_t shift,
                   uint32_t* dir, void* stream) {
  SQDK_REQUIRE(keys && dir, "NULL pointer argument");
  SQDK_REQUIRE(n >= 1 && n < (1ll << 31), "n=%lld out of range", (long long)n);
  SQDK_REQUIRE(shift >= 0 && shift <= num_qubits && shift <= 63 && num_qubits - shift <= 26,
               "shift %d invalid for %d qubits", shift, num_qubits);
  const int64_t np = 1ll << (num_qubits - shift);
  dir_kernel<<<(unsigned)ceil_div(np + 1, kBlock), kBlock, 0, as_stream(stream)>>>(
      keys, n, shift, np, dir);
  SQDK_LAUNCHED("dir_kernel");
  return SQDK_OK;
}

int sqdk_space_fine(const uint64_t* keys, int64_t n, int32_t num_qubits, int32_t bucket_bits,
                    int32_t fine_bits, const uint32_t* bstart, uint16_t* fdir, void* stream) {
  SQDK_REQUIRE(keys && bstart && fdir, "NULL pointer argument");
  SQDK_REQUIRE(n >= 1 && n < (1ll << 31), "n=%lld out of range", (long long)n);
  SQDK_REQUIRE(bucket_bits >= 0 && bucket_bits <= num_qubits && bucket_bits <= 63 &&
                   num_qubits - bucket_bits <= 26,
               "bucket_bits %d invalid", bucket_bits);
  SQDK_REQUIRE(fine_bits >= 0 && fine_bits <= bucket_bits && fine_bits <= 15,
               "fine_bits %d invalid", fine_bits);
  const int64_t total = (1ll << (num_qubits - bucket_bits)) << fine_bits;
  SQDK_REQUIRE(total <= kMaxFineEntries, "fine directory of %lld entries (limit 2^31)",
               (long long)total);
  fine_kernel<<<grid_stride_blocks(total), kBlock, 0, as_stream(stream)>>>(
      keys, bstart, bucket_bits, fine_bits, total, fdir);
  SQDK_LAUNCHED("fine_kernel");
  return SQDK_OK;
}

int sqdk_space_bits(const uint64_t* keys, int64_t n, int32_t num_qubits, int32_t bucket_bits,
                    int32_t fine_bits, const uint32_t* bstart, const uint16_t* fdir,
                    uint32_t* fbits, void* stream) {
  SQDK_REQUIRE(keys && bstart && fdir && fbits, "NULL pointer argument");
  SQDK_REQUIRE(n >= 1 && n < (1ll << 31), "n=%lld out of range", (long long)n);
  SQDK_REQUIRE(bucket_bits >= 5 && bucket_bits <= num_qubits && bucket_bits <= 20 &&
                   num_qubits - bucket_bits <= 26,
               "bucket_bits %d invalid for the dense lookup", bucket_bits);
  SQDK_REQUIRE(fine_bits == bucket_bits - 5, "dense lookup needs fine_bits = bucket_bits - 5 "
               "(got %d, %d)", fine_bits, bucket_bits);
  const int64_t total = (1ll << (num_qubits - bucket_bits)) << fine_bits;
  SQDK_REQUIRE(total <= kMaxFineEntries, "fine directory of %lld entries (limit 2^31)",
               (long long)total);
  bits_kernel<<<grid_stride_blocks(total), kBlock, 0, as_stream(stream)>>>(
      keys, bstart, fdir, fine_bits, total, fbits);
  SQDK_LAUNCHED("bits_kernel");
  return SQDK_OK;
}

int sqdk_space_items(const uint32_t* bstart, int64_t n_buckets, int32_t chunk_rows,
                     uint32_t* items, int64_t max_items, int32_t* n_items_out, void* stream) {
  SQDK_REQUIRE(bstart && items && n_items_out, "NULL pointer argument");
  SQDK_REQUIRE(n_buckets >= 1 && chunk_rows >= 1, "bad bucket count or chunk size");
  items_kernel<<<1, kBlock, 0, as_stream(stream)>>>(bstart, n_buckets, chunk_rows, items,
                                                     max_items, n_items_out);
  SQDK_LAUNCHED("items_kernel");
  return SQDK_OK;
}

int sqdk_xsource(const sqdk_space* space, uint64_t xmask, int32_t* out, void* stream) {
  SQDK_TRY(check_space(space));
  SQDK_REQUIRE(out, "NULL output");
  xsource_kernel<<<(unsigned)ceil_div(space->n, kBlock), kBlock, 0, as_stream(stream)>>>(
      space->keys, space->bstart, space->n, space->num_qubits, space->bucket_bits, xmask, out);
  SQDK_LAUNCHED("xsource_kernel");
  return SQDK_OK;
}

int sqdk_diagonal(const uint64_t* zmask, const double* coef, int32_t t0, int32_t t1,
                  const uint64_t* keys, int64_t n, double* out, void* stream) {
  SQDK_REQUIRE(keys && out, "NULL pointer argument");
  SQDK_REQUIRE(t0 >= 0 && t1 >= t0, "bad term range [%d, %d)", t0, t1);
  SQDK_REQUIRE(t1 == t0 || (zmask && coef), "NULL term arrays");
  SQDK_REQUIRE(n >= 1 && n < (1ll << 31), "n=%lld out of range", (long long)n);
  diagonal_kernel<<<(unsigned)ceil_div(n, kBlock), kBlock, 0, as_stream(stream)>>>(
      zmask, reinterpret_cast<const double2*>(coef), t0, t1, keys, n,
      reinterpret_cast<double2*>(out));
  SQDK_LAUNCHED("diagonal_kernel");
  return SQDK_OK;
}

}  // extern "C"
