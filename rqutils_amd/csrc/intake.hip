// Intake of sampled bitstrings on the GPU: pack rows into uint64 keys, sort
// the keys with a stable LSD radix sort, drop duplicates, and unpack keys back
// into rows. Counterpart of the reference's row packing and uniquify_states
// (rqutils/sqd.py:237, 414-445: a lexsort of packed rows and a unique with
// fill rows) and of the basis-state unpack (sqd.py:246).
//
// Key order: bit q of a key is qubit q, and column c of a state row is qubit
// n-1-c, so ascending keys are the reference's lexicographic order of packed
// rows. The sort runs ceil(key_bits / 4) passes of 4-bit digits. Each pass is
// a per-tile digit histogram, one exclusive scan of the histograms in digit-
// major order (every (digit, tile) gets its global start) and a scatter in
// which every key's rank within its tile is counted per thread in LDS and then
// offset by a block scan, so keys of equal digit keep their input order.
// Unique is a per-tile count of "differs from the previous key", a scan of the
// tile counts and a compaction. Everything is integer work: bit-exact.
#include "internal.hpp"

namespace sqdk {

namespace {

constexpr int kSortThreads = 256;
constexpr int kSortItems = 16;                       // keys per thread, consecutive
constexpr int kTile = kSortThreads * kSortItems;     // keys per workgroup
constexpr int kDigitBits = 4;
constexpr int kDigits = 1 << kDigitBits;
constexpr int kScanThreads = 1024;
static_assert(kDigits * kSortThreads == kSortThreads * kSortItems,
              "the rank scan gives each thread kDigits consecutive counters");

// Row c of a state (column c) is qubit n-1-c; any nonzero byte is a 1.
__global__ void __launch_bounds__(256)
pack_kernel(const uint8_t* __restrict__ states, int64_t n_rows, int nq, int64_t stride,
            uint64_t* __restrict__ keys) {
  const int64_t row = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (row >= n_rows) return;
  const uint8_t* p = states + row * stride;
  uint64_t k = 0;
  for (int c = 0; c < nq; ++c) k |= (uint64_t)(p[c] != 0) << (nq - 1 - c);
  keys[row] = k;
}

__global__ void __launch_bounds__(256)
unpack_kernel(const uint64_t* __restrict__ keys, int64_t n, int nq, uint8_t* __restrict__ out) {
  const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= n * nq) return;
  const int64_t row = idx / nq;
  const int c = (int)(idx % nq);
  out[idx] = (uint8_t)((keys[row] >> (nq - 1 - c)) & 1ull);
}

// Exclusive scan of one value per thread over the workgroup; every thread must
// call it. ws holds blockDim.x / 64 + 1 words. Returns the thread's prefix
// and sets total.
__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t* ws, uint32_t& total) {
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int nw = blockDim.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t y = __shfl_up(x, off);
    if (lane >= off) x += y;
  }
  if (lane == 63) ws[w] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t s = 0;
    for (int i = 0; i < nw; ++i) {
      const uint32_t t = ws[i];
      ws[i] = s;
      s += t;
    }
    ws[nw] = s;
  }
  __syncthreads();
  total = ws[nw];
  const uint32_t r = ws[w] + x - v;
  __syncthreads();  // ws may be reused after the return
  return r;
}

// hist[d * tiles + t] = keys of tile t whose digit at `shift` is d.
__global__ void __launch_bounds__(kSortThreads)
radix_hist_kernel(const uint64_t* __restrict__ keys, int64_t n, int shift, int64_t tiles,
                  uint32_t* __restrict__ hist) {
  __shared__ uint32_t cnt[kDigits];
  const int tid = threadIdx.x;
  if (tid < kDigits) cnt[tid] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kTile;
  for (int j = 0; j < kSortItems; ++j) {
    const int64_t i = base + (int64_t)j * kSortThreads + tid;
    if (i < n) atomicAdd(&cnt[(uint32_t)(keys[i] >> shift) & (kDigits - 1)], 1u);
  }
  __syncthreads();
  if (tid < kDigits) hist[(int64_t)tid * tiles + blockIdx.x] = cnt[tid];
}

// In-place exclusive scan of a[0..m) by one workgroup; *total = sum.
__global__ void __launch_bounds__(kScanThreads)
exclusive_scan_kernel(uint32_t* __restrict__ a, int64_t m, uint32_t* __restrict__ total) {
  __shared__ uint32_t ws[kScanThreads / 64 + 1];
  const int64_t per = (m + kScanThreads - 1) / kScanThreads;
  const int64_t lo = min(m, (int64_t)threadIdx.x * per);
  const int64_t hi = min(m, lo + per);
  uint32_t s = 0;
  for (int64_t i = lo; i < hi; ++i) s += a[i];
  uint32_t tot;
  uint32_t run = block_exclusive_scan(s, ws, tot);
  for (int64_t i = lo; i < hi; ++i) {
    const uint32_t t = a[i];
    a[i] = run;
    run += t;
  }
  if (total && threadIdx.x == 0) *total = tot;
}

// Stable scatter of one pass: offs = the scanned histograms.
__global__ void __launch_bounds__(kSortThreads)
radix_scatter_kernel(const uint64_t* __restrict__ in, uint64_t* __restrict__ out, int64_t n,
                     int shift, int64_t tiles, const uint32_t* __restrict__ offs) {
  __shared__ uint32_t cnt[kDigits * kSortThreads];  // [digit][thread]
  __shared__ uint32_t ws[kSortThreads / 64 + 1];
  const int tid = threadIdx.x;
#pragma unroll
  for (int d = 0; d < kDigits; ++d) cnt[d * kSortThreads + tid] = 0;
  // The thread's keys are consecutive, so counting them in order keeps equal
  // digits in input order within the thread.
  const int64_t base = (int64_t)blockIdx.x * kTile + (int64_t)tid * kSortItems;
  uint64_t k[kSortItems];
  uint32_t rank[kSortItems];
#pragma unroll
  for (int j = 0; j < kSortItems; ++j) {
    const int64_t i = base + j;
    k[j] = i < n ? in[i] : 0ull;
    rank[j] = 0;
    if (i < n) {
      const int d = (int)((uint32_t)(k[j] >> shift) & (kDigits - 1));
      rank[j] = cnt[d * kSortThreads + tid]++;
    }
  }
  __syncthreads();
  // Digit-major exclusive scan of the counters: thread u owns entries
  // u*kDigits .. u*kDigits + kDigits - 1 of the [digit][thread] array.
  uint32_t mine[kDigits];
  uint32_t s = 0;
#pragma unroll
  for (int e = 0; e < kDigits; ++e) {
    mine[e] = cnt[tid * kDigits + e];
    s += mine[e];
  }
  uint32_t tot;
  uint32_t run = block_exclusive_scan(s, ws, tot);
#pragma unroll
  for (int e = 0; e < kDigits; ++e) {
    cnt[tid * kDigits + e] = run;
    run += mine[e];
  }
  __syncthreads();
  // cnt[d][t]: keys of the tile with a smaller digit, or digit d in an earlier
  // thread; cnt[d][0] is where digit d starts in the tile.
#pragma unroll
  for (int j = 0; j < kSortItems; ++j) {
    const int64_t i = base + j;
    if (i < n) {
      const int d = (int)((uint32_t)(k[j] >> shift) & (kDigits - 1));
      const uint32_t pos = cnt[d * kSortThreads + tid] + rank[j] - cnt[d * kSortThreads];
      out[(int64_t)offs[(int64_t)d * tiles + blockIdx.x] + pos] = k[j];
    }
  }
}

__device__ inline bool first_of_run(const uint64_t* __restrict__ k, int64_t i) {
  return i == 0 || k[i] != k[i - 1];
}

// tile_cnt[t] = distinct-key starts in tile t of the sorted keys.
__global__ void __launch_bounds__(kSortThreads)
unique_count_kernel(const uint64_t* __restrict__ k, int64_t n, uint32_t* __restrict__ tile_cnt) {
  __shared__ uint32_t ws[kSortThreads / 64 + 1];
  const int64_t base = (int64_t)blockIdx.x * kTile + (int64_t)threadIdx.x * kSortItems;
  uint32_t s = 0;
  for (int j = 0; j < kSortItems; ++j) {
    const int64_t i = base + j;
    if (i < n && first_of_run(k, i)) ++s;
  }
  uint32_t tot;
  block_exclusive_scan(s, ws, tot);
  if (threadIdx.x == 0) tile_cnt[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(kSortThreads)
unique_write_kernel(const uint64_t* __restrict__ k, int64_t n,
                    const uint32_t* __restrict__ tile_off, uint64_t* __restrict__ out) {
  __shared__ uint32_t ws[kSortThreads / 64 + 1];
  const int64_t base = (int64_t)blockIdx.x * kTile + (int64_t)threadIdx.x * kSortItems;
  uint32_t s = 0;
  for (int j = 0; j < kSortItems; ++j) {
    const int64_t i = base + j;
    if (i < n && first_of_run(k, i)) ++s;
  }
  uint32_t tot;
  uint32_t pos = tile_off[blockIdx.x] + block_exclusive_scan(s, ws, tot);
  for (int j = 0; j < kSortItems; ++j) {
    const int64_t i = base + j;
    if (i < n && first_of_run(k, i)) out[pos++] = k[i];
  }
}

__global__ void store_count_kernel(const uint32_t* __restrict__ total, int64_t* __restrict__ out) {
  if (threadIdx.x == 0) *out = (int64_t)*total;
}

struct IntakeWork {
  uint64_t* a;
  uint64_t* b;
  uint32_t* hist;
  uint32_t* tile_cnt;
  uint32_t* total;
  size_t bytes;
};

inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

IntakeWork intake_layout(int64_t n, void* base) {
  const size_t tiles = (size_t)ceil_div(n, kTile);
  const size_t off_b = align256(sizeof(uint64_t) * (size_t)n);
  const size_t off_hist = off_b + align256(sizeof(uint64_t) * (size_t)n);
  const size_t off_cnt = off_hist + align256(sizeof(uint32_t) * kDigits * tiles);
  const size_t off_total = off_cnt + align256(sizeof(uint32_t) * tiles);
  IntakeWork w;
  w.bytes = off_total + 256;
  char* p = static_cast<char*>(base);
  w.a = p ? reinterpret_cast<uint64_t*>(p) : nullptr;
  w.b = p ? reinterpret_cast<uint64_t*>(p + off_b) : nullptr;
  w.hist = p ? reinterpret_cast<uint32_t*>(p + off_hist) : nullptr;
  w.tile_cnt = p ? reinterpret_cast<uint32_t*>(p + off_cnt) : nullptr;
  w.total = p ? reinterpret_cast<uint32_t*>(p + off_total) : nullptr;
  return w;
}

}  // namespace

}  // namespace sqdk

using namespace sqdk;

extern "C" {

int sqdk_intake_pack(const uint8_t* states, int64_t n_rows, int32_t num_qubits,
                     int64_t row_stride, uint64_t* keys, void* stream) {
  SQDK_REQUIRE(states && keys, "NULL pointer argument");
  SQDK_REQUIRE(n_rows >= 1 && n_rows < (1ll << 31), "n_rows=%lld out of range",
               (long long)n_rows);
  SQDK_REQUIRE(num_qubits >= 1 && num_qubits <= 64, "num_qubits %d out of range", num_qubits);
  SQDK_REQUIRE(row_stride >= num_qubits, "row stride %lld below %d columns",
               (long long)row_stride, num_qubits);
  pack_kernel<<<(unsigned)ceil_div(n_rows, 256), 256, 0, as_stream(stream)>>>(
      states, n_rows, num_qubits, row_stride, keys);
  SQDK_LAUNCHED("pack_kernel");
  return SQDK_OK;
}

int sqdk_intake_unique_workspace(int64_t n, size_t* bytes) {
  SQDK_REQUIRE(bytes, "NULL output");
  SQDK_REQUIRE(n >= 1 && n < (1ll << 31), "n=%lld out of range", (long long)n);
  *bytes = intake_layout(n, nullptr).bytes;
  return SQDK_OK;
}

int sqdk_intake_unique(const uint64_t* keys, int64_t n, int32_t key_bits, uint64_t* out,
                       int64_t* n_unique, void* work, size_t work_bytes, void* stream) {
  SQDK_REQUIRE(keys && out && n_unique && work, "NULL pointer argument");
  SQDK_REQUIRE(n >= 1 && n < (1ll << 31), "n=%lld out of range", (long long)n);
  SQDK_REQUIRE(key_bits >= 1 && key_bits <= 64, "key_bits %d out of range", key_bits);
  const IntakeWork w = intake_layout(n, work);
  SQDK_REQUIRE(work_bytes >= w.bytes, "intake workspace of %zu bytes below %zu", work_bytes,
               w.bytes);
  hipStream_t s = as_stream(stream);
  const int64_t tiles = ceil_div(n, kTile);
  SQDK_REQUIRE(tiles < (1ll << 31), "too many tiles");
  const int passes = (key_bits + kDigitBits - 1) / kDigitBits;
  uint64_t* bufs[2] = {w.a, w.b};
  for (int p = 0; p < passes; ++p) {
    const uint64_t* src = p == 0 ? keys : bufs[(p + 1) & 1];
    uint64_t* dst = bufs[p & 1];
    const int shift = p * kDigitBits;
    radix_hist_kernel<<<(unsigned)tiles, kSortThreads, 0, s>>>(src, n, shift, tiles, w.hist);
    SQDK_LAUNCHED("radix_hist_kernel");
    exclusive_scan_kernel<<<1, kScanThreads, 0, s>>>(w.hist, (int64_t)kDigits * tiles, nullptr);
    SQDK_LAUNCHED("exclusive_scan_kernel");
    radix_scatter_kernel<<<(unsigned)tiles, kSortThreads, 0, s>>>(src, dst, n, shift, tiles,
                                                                 w.hist);
    SQDK_LAUNCHED("radix_scatter_kernel");
  }
  const uint64_t* sorted = bufs[(passes - 1) & 1];
  unique_count_kernel<<<(unsigned)tiles, kSortThreads, 0, s>>>(sorted, n, w.tile_cnt);
  SQDK_LAUNCHED("unique_count_kernel");
  exclusive_scan_kernel<<<1, kScanThreads, 0, s>>>(w.tile_cnt, tiles, w.total);
  SQDK_LAUNCHED("exclusive_scan_kernel");
  unique_write_kernel<<<(unsigned)tiles, kSortThreads, 0, s>>>(sorted, n, w.tile_cnt, out);
  SQDK_LAUNCHED("unique_write_kernel");
  store_count_kernel<<<1, 64, 0, s>>>(w.total, n_unique);
  SQDK_LAUNCHED("store_count_kernel");
  return SQDK_OK;
}

int sqdk_intake_unpack(const uint64_t* keys, int64_t n, int32_t num_qubits, uint8_t* states,
                       void* stream) {
  SQDK_REQUIRE(keys && states, "NULL pointer argument");
  SQDK_REQUIRE(n >= 1 && n < (1ll << 31), "n=%lld out of range", (long long)n);
  SQDK_REQUIRE(num_qubits >= 1 && num_qubits <= 64, "num_qubits %d out of range", num_qubits);
  unpack_kernel<<<(unsigned)ceil_div(n * num_qubits, 256), 256, 0, as_stream(stream)>>>(
      keys, n, num_qubits, states);
  SQDK_LAUNCHED("unpack_kernel");
  return SQDK_OK;
}

}  // extern "C"
