// Intake of sampled bitstrings on the GPU: pack rows into uint64 keys, sort
// the keys and drop duplicates, and unpack keys back into rows. Counterpart of
// the reference's row packing and uniquify_states (rqutils/sqd.py:237,
// 414-445: a lexsort of packed rows and a unique with fill rows) and of the
// basis-state unpack (sqd.py:246).
//
// Key order: bit q of a key is qubit q, and column c of a state row is qubit
// n-1-c, so ascending keys are the reference's lexicographic order of packed
// rows. Two ways to the sorted distinct keys, both bit-exact integer work:
//
// * Occupancy bitmap, for keys of at most 28 bits (the rule is bitmap_fits).
//   The key space is cut into slices of 2^20 keys, whose bitmap (128 KiB)
//   fits in LDS. One stable 8-bit counting pass groups the keys by slice;
//   then a workgroup per (slice, share of its keys) sets the bits of its keys
//   with LDS atomics and stores the slice's bitmap. A popcount per 4,096-word
//   block (OR-ing the shares' bitmaps), a scan of the block counts and one
//   pass that writes the position of every set bit give the distinct keys in
//   ascending order. No global atomics: on this GPU those run at a small
//   fraction of HBM speed.
// * Stable LSD radix sort otherwise: ceil(key_bits / 8) passes of 8-bit
//   digits. Each pass is a per-tile digit histogram, an exclusive scan of each
//   digit's histogram over the tiles, and a scatter that ranks every key of a
//   4,096-key tile among equal digits with wave ballots (so equal digits keep
//   their input order), orders the tile in LDS and writes each digit's run
//   contiguously. Unique is a per-tile count of "differs from the previous
//   key", a scan of the tile counts and a compaction.
//
// Keys not below 2^key_bits are an input error: both ways count them and
// report n_unique = -1 instead of a result.
#include "internal.hpp"

namespace sqdk {

namespace {

constexpr int kSortThreads = 256;
constexpr int kSortItems = 16;                       // keys per thread
constexpr int kTile = kSortThreads * kSortItems;     // keys per workgroup
constexpr int kDigitBits = 8;
constexpr int kDigits = 1 << kDigitBits;
constexpr int kSortWaves = kSortThreads / 64;
constexpr int kScanThreads = 1024;
static_assert(kDigits == kSortThreads, "the scatter gives each thread one digit");
static_assert(kTile == kSortWaves * 64 * kSortItems, "a wave takes 64 * kSortItems keys");

constexpr int kRowThreads = 256;                     // rows per pack / unpack workgroup

// Row c of a state (column c) is qubit n-1-c; any nonzero byte is a 1. A
// workgroup packs 256 rows. Dense rows (stride == nq) of a 4-byte aligned
// array are one run of 256 * nq bytes starting on a 4-byte boundary: it is
// staged through LDS with coalesced 4-byte loads, then each thread reads its
// row from LDS. Other layouts read each row's bytes directly.
__global__ void __launch_bounds__(kRowThreads)
pack_kernel(const uint8_t* __restrict__ states, int64_t n_rows, int nq, int64_t stride,
            uint64_t* __restrict__ keys) {
  __shared__ uint32_t tile[kRowThreads * 64 / 4];   // 256 rows of at most 64 bytes
  const int tid = threadIdx.x;
  const int64_t row0 = (int64_t)blockIdx.x * kRowThreads;
  const int rows = (int)min<int64_t>(kRowThreads, n_rows - row0);
  const uint8_t* p;
  if (stride == nq && (reinterpret_cast<uintptr_t>(states) & 3) == 0) {
    const uint8_t* src = states + row0 * nq;
    const int bytes = rows * nq;
    const int words = bytes >> 2;
    const uint32_t* src4 = reinterpret_cast<const uint32_t*>(src);
    // at most 16 words per thread; all loads are issued before the LDS stores
    uint32_t v[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int i = r * kRowThreads + tid;
      v[r] = i < words ? src4[i] : 0u;
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int i = r * kRowThreads + tid;
      if (i < words) tile[i] = v[r];
    }
    uint8_t* tb = reinterpret_cast<uint8_t*>(tile);
    if (tid < (bytes & 3)) tb[words * 4 + tid] = src[words * 4 + tid];
    __syncthreads();
    p = tb + tid * nq;
  } else {
    p = states + (row0 + tid) * stride;
  }
  if (tid >= rows) return;
  uint64_t k = 0;
#pragma unroll 4
  for (int c = 0; c < nq; ++c) k |= (uint64_t)(p[c] != 0) << (nq - 1 - c);
  keys[row0 + tid] = k;
}

// The inverse: a workgroup writes 256 rows, built in LDS and stored as one run
// of 256 * nq bytes (4-byte stores when the output is 4-byte aligned).
__global__ void __launch_bounds__(kRowThreads)
unpack_kernel(const uint64_t* __restrict__ keys, int64_t n, int nq, uint8_t* __restrict__ out) {
  __shared__ uint32_t tile[kRowThreads * 64 / 4];
  uint8_t* tb = reinterpret_cast<uint8_t*>(tile);
  const int tid = threadIdx.x;
  const int64_t row0 = (int64_t)blockIdx.x * kRowThreads;
  const int rows = (int)min<int64_t>(kRowThreads, n - row0);
  if (tid < rows) {
    const uint64_t k = keys[row0 + tid];
    for (int c = 0; c < nq; ++c) tb[tid * nq + c] = (uint8_t)((k >> (nq - 1 - c)) & 1ull);
  }
  __syncthreads();
  uint8_t* dst = out + row0 * nq;
  const int bytes = rows * nq;
  if ((reinterpret_cast<uintptr_t>(out) & 3) == 0) {
    const int words = bytes >> 2;
    uint32_t* dst4 = reinterpret_cast<uint32_t*>(dst);
    for (int i = tid; i < words; i += kRowThreads) dst4[i] = tile[i];
    if (tid < (bytes & 3)) dst[words * 4 + tid] = tb[words * 4 + tid];
  } else {
    for (int i = tid; i < bytes; i += kRowThreads) dst[i] = tb[i];
  }
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

// hist[d * tiles + t] = keys of tile t (keys [t * 4096, (t + 1) * 4096))
// whose digit (k >> shift) & dmask is d, for d <= dmask. With `bad` given,
// keys not below 2^key_bits are also counted into *bad.
__global__ void __launch_bounds__(kSortThreads)
radix_hist_kernel(const uint64_t* __restrict__ keys, int64_t n, int shift, uint32_t dmask,
                  int64_t tiles, uint32_t* __restrict__ hist, int key_bits,
                  uint32_t* __restrict__ bad) {
  __shared__ uint32_t cnt[kSortWaves][kDigits];      // one histogram per wave
  const int tid = threadIdx.x;
  const int w = tid >> 6;
#pragma unroll
  for (int v = 0; v < kSortWaves; ++v) cnt[v][tid] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kTile;
  bool out_of_range = false;
  uint64_t k[kSortItems];
#pragma unroll
  for (int j = 0; j < kSortItems; ++j)
    k[j] = keys[min(base + (int64_t)j * kSortThreads + tid, n - 1)];
#pragma unroll
  for (int j = 0; j < kSortItems; ++j) {
    if (base + (int64_t)j * kSortThreads + tid < n) {
      out_of_range |= key_bits < 64 && (k[j] >> key_bits) != 0;
      atomicAdd(&cnt[w][(uint32_t)(k[j] >> shift) & dmask], 1u);
    }
  }
  if (bad && out_of_range) atomicAdd(bad, 1u);
  __syncthreads();
  if ((uint32_t)tid <= dmask) {
    uint32_t c = 0;
#pragma unroll
    for (int v = 0; v < kSortWaves; ++v) c += cnt[v][tid];
    hist[(int64_t)tid * tiles + blockIdx.x] = c;
  }
}

// In-place exclusive scan of a[b*m .. (b+1)*m) by workgroup b; total[b] = its
// sum (when total is given).
__global__ void __launch_bounds__(kScanThreads)
exclusive_scan_kernel(uint32_t* __restrict__ a, int64_t m, uint32_t* __restrict__ total) {
  __shared__ uint32_t ws[kScanThreads / 64 + 1];
  a += (int64_t)blockIdx.x * m;
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
  if (total && threadIdx.x == 0) total[blockIdx.x] = tot;
}

// Stable scatter of one pass. offs[d * tiles + t] = keys of digit d in tiles
// before t (each digit's histogram scanned over the tiles); dtot[d] = keys of
// digit d in all. Wave w of tile t takes keys t * 4096 + w * 1024 + j * 64 +
// lane, j = 0..15, in that (input) order: the lanes holding the same digit are
// found with dbits ballots, and a per-wave LDS counter per digit gives each
// key its rank among equal digits before it. A scan over digits and waves
// turns the ranks into positions of the sorted tile, which is assembled in
// LDS and written out one digit run after another.
__global__ void __launch_bounds__(kSortThreads)
radix_scatter_kernel(const uint64_t* __restrict__ in, uint64_t* __restrict__ out, int64_t n,
                     int shift, int dbits, int64_t tiles, const uint32_t* __restrict__ offs,
                     const uint32_t* __restrict__ dtot) {
  __shared__ uint64_t stage[kTile];
  __shared__ uint32_t cnt[kSortWaves][kDigits];
  __shared__ uint32_t gbase[kDigits];
  __shared__ uint32_t ws[kSortWaves + 1];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int w = tid >> 6;
  const uint32_t dmask = (1u << dbits) - 1u;
#pragma unroll
  for (int v = 0; v < kSortWaves; ++v) cnt[v][tid] = 0;
  __syncthreads();
  const int64_t tile0 = (int64_t)blockIdx.x * kTile;
  const int64_t base = tile0 + (int64_t)w * (64 * kSortItems);
  uint64_t k[kSortItems];
  uint32_t rank[kSortItems];
  // All 16 loads first (indices clamped into the array, so none is
  // conditional); lanes past n are masked out of the ranking below.
#pragma unroll
  for (int j = 0; j < kSortItems; ++j) k[j] = in[min(base + j * 64 + lane, n - 1)];
#pragma unroll
  for (int j = 0; j < kSortItems; ++j) {
    const bool valid = base + j * 64 + lane < n;
    const uint32_t d = (uint32_t)(k[j] >> shift) & dmask;
    uint64_t peers = __ballot(valid);
    for (int b = 0; b < dbits; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t bal = __ballot(bit);
      peers &= bit ? bal : ~bal;
    }
    const uint32_t before = __builtin_amdgcn_mbcnt_hi(
        (uint32_t)(peers >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)peers, 0u));
    const uint32_t old = cnt[w][d];
    rank[j] = old + before;
    // The lowest lane of each digit group advances the wave's counter after
    // every lane of the group has read it (one wave: LDS ops stay in order).
    if (valid && before == 0) cnt[w][d] = old + (uint32_t)__popcll(peers);
  }
  __syncthreads();
  // Thread d: start of digit d in the sorted tile, then of each wave's keys
  // of digit d; the global start of the digit's run for this tile.
  {
    const int d = tid;
    uint32_t c[kSortWaves];
    uint32_t tot = 0;
#pragma unroll
    for (int v = 0; v < kSortWaves; ++v) {
      c[v] = cnt[v][d];
      tot += c[v];
    }
    uint32_t unused;
    const uint32_t start = block_exclusive_scan(tot, ws, unused);
    const uint32_t dsum = (uint32_t)d <= dmask ? dtot[d] : 0u;
    const uint32_t gstart = block_exclusive_scan(dsum, ws, unused);
    uint32_t run = start;
#pragma unroll
    for (int v = 0; v < kSortWaves; ++v) {
      cnt[v][d] = run;
      run += c[v];
    }
    if ((uint32_t)d <= dmask)
      gbase[d] = gstart + offs[(int64_t)d * tiles + blockIdx.x] - start;
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kSortItems; ++j) {
    const int64_t i = base + j * 64 + lane;
    if (i < n) {
      const uint32_t d = (uint32_t)(k[j] >> shift) & dmask;
      stage[cnt[w][d] + rank[j]] = k[j];
    }
  }
  __syncthreads();
  const int count = (int)min<int64_t>(kTile, n - tile0);
  for (int e = tid; e < count; e += kSortThreads) {
    const uint64_t key = stage[e];
    const uint32_t d = (uint32_t)(key >> shift) & dmask;
    out[gbase[d] + (uint32_t)e] = key;   // 32-bit sum: gbase may wrap, the position does not
  }
}

// Unique over sorted keys, with the scatter's wave-striped layout: wave w of
// tile t takes keys t * 4096 + w * 1024 + j * 64 + lane, j = 0..15 (coalesced).
// A key starts a run when it is the first or differs from its predecessor.
__device__ inline void run_starts(const uint64_t* __restrict__ k, int64_t n, int64_t base,
                                 int lane, uint64_t (&key)[kSortItems],
                                 bool (&first)[kSortItems]) {
  uint64_t prev[kSortItems];
#pragma unroll
  for (int j = 0; j < kSortItems; ++j) {
    const int64_t i = min(base + j * 64 + lane, n - 1);
    const int64_t ip = i > 0 ? i - 1 : 0;
    key[j] = k[i];
    prev[j] = k[ip];
  }
#pragma unroll
  for (int j = 0; j < kSortItems; ++j) {
    const int64_t i = base + j * 64 + lane;
    first[j] = i < n && (i == 0 || key[j] != prev[j]);
  }
}

// tile_cnt[t] = run starts in tile t of the sorted keys.
__global__ void __launch_bounds__(kSortThreads)
unique_count_kernel(const uint64_t* __restrict__ k, int64_t n, uint32_t* __restrict__ tile_cnt) {
  __shared__ uint32_t ws[kSortWaves + 1];
  const int lane = threadIdx.x & 63;
  const int64_t base = (int64_t)blockIdx.x * kTile + (int64_t)(threadIdx.x >> 6) * (64 * kSortItems);
  uint64_t key[kSortItems];
  bool first[kSortItems];
  run_starts(k, n, base, lane, key, first);
  uint32_t s = 0;
#pragma unroll
  for (int j = 0; j < kSortItems; ++j) s += first[j] ? 1u : 0u;
  uint32_t tot;
  block_exclusive_scan(s, ws, tot);
  if (threadIdx.x == 0) tile_cnt[blockIdx.x] = tot;
}

// Writes the run starts of tile t from out[tile_off[t]] on, in order: wave
// totals are scanned over the tile, then each step's starts are placed by a
// ballot and mbcnt.
__global__ void __launch_bounds__(kSortThreads)
unique_write_kernel(const uint64_t* __restrict__ k, int64_t n,
                    const uint32_t* __restrict__ tile_off, uint64_t* __restrict__ out) {
  __shared__ uint32_t wsum[kSortWaves];
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int64_t base = (int64_t)blockIdx.x * kTile + (int64_t)w * (64 * kSortItems);
  uint64_t key[kSortItems];
  bool first[kSortItems];
  run_starts(k, n, base, lane, key, first);
  uint64_t bal[kSortItems];
  uint32_t t = 0;
#pragma unroll
  for (int j = 0; j < kSortItems; ++j) {
    bal[j] = __ballot(first[j]);
    t += (uint32_t)__popcll(bal[j]);
  }
  if (lane == 0) wsum[w] = t;
  __syncthreads();
  uint32_t pos = tile_off[blockIdx.x];
  for (int v = 0; v < w; ++v) pos += wsum[v];
#pragma unroll
  for (int j = 0; j < kSortItems; ++j) {
    const uint32_t below = __builtin_amdgcn_mbcnt_hi(
        (uint32_t)(bal[j] >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bal[j], 0u));
    if (first[j]) out[pos + below] = key[j];
    pos += (uint32_t)__popcll(bal[j]);
  }
}

// Occupancy bitmap of keys below 2^key_bits: bit (k & 31) of word k >> 5.
constexpr int kBmThreads = 256;
constexpr int kBmLaneWords = 16;                             // words per lane
constexpr int kBmWaveWords = 64 * kBmLaneWords;              // 1,024 words per wave
constexpr int kBmBlockWords = kBmWaveWords * (kBmThreads / 64);  // 4,096 per workgroup
constexpr int kBitmapMaxBits = 28;
constexpr int kSliceBits = 20;                               // 2^20 keys: 128 KiB of LDS
constexpr int kSliceThreads = 1024;
constexpr int kSliceBlocks = 512;                            // target grid of the slice pass

inline int64_t bitmap_words(int key_bits) {
  return key_bits > 5 ? (int64_t)1 << (key_bits - 5) : 1;
}

// The bitmap way for keys of at most 28 bits whose bitmap is small or at
// most 32 bytes per key (its passes over 2^key_bits / 8 bytes then cost about
// one radix pass or less).
inline bool bitmap_fits(int64_t n, int key_bits) {
  if (key_bits > kBitmapMaxBits) return false;
  const int64_t bytes = bitmap_words(key_bits) * 4;
  return bytes <= ((int64_t)4 << 20) || bytes <= 32 * n;
}

// How the bitmap way splits the work: 2^pbits slices of 2^sbits keys, each
// built by `shares` workgroups (which then leave one bitmap each).
struct BitmapPlan {
  int sbits, pbits;
  int64_t slices, shares, words, slice_words;
};

inline BitmapPlan bitmap_plan(int64_t n, int key_bits) {
  BitmapPlan b;
  b.sbits = std::min(key_bits, kSliceBits);
  b.pbits = key_bits - b.sbits;
  b.slices = (int64_t)1 << b.pbits;
  b.shares = std::max<int64_t>(1, std::min(ceil_div(kSliceBlocks, b.slices),
                                           ceil_div(n, b.slices * kTile)));
  b.words = bitmap_words(key_bits);
  b.slice_words = bitmap_words(b.sbits);
  return b;
}

// Workgroup (slice p, share c): the keys of slice p are part[start_p, end_p)
// (start_p = keys of earlier slices per dtot; all n keys when there is one
// slice); share c takes every shares-th run of 512 of them, sets their bits in
// an LDS bitmap and stores it at bits + c * words + p * slice_words. Keys not
// below 2^key_bits set nothing and are counted into *bad.
__global__ void __launch_bounds__(kSliceThreads)
bitmap_slice_kernel(const uint64_t* __restrict__ part, int64_t n, int key_bits, int sbits,
                    int64_t slices, int64_t shares, const uint32_t* __restrict__ dtot,
                    int64_t words, int64_t slice_words, uint32_t* __restrict__ bits,
                    uint32_t* __restrict__ bad) {
  __shared__ uint32_t bm[1 << (kSliceBits - 5)];
  __shared__ uint32_t ws[kSliceThreads / 64 + 1];
  const int tid = threadIdx.x;
  const int64_t p = blockIdx.x / shares;
  const int64_t c = blockIdx.x % shares;
  for (int64_t e = tid; e < slice_words; e += kSliceThreads) bm[e] = 0;
  int64_t start = 0, end = n;
  if (slices > 1) {
    // slices <= 256: thread q < p adds dtot[q]
    uint32_t tot;
    block_exclusive_scan(tid < p ? dtot[tid] : 0u, ws, tot);
    start = tot;
    end = start + dtot[p];
  } else {
    __syncthreads();
  }
  const uint64_t lmask = ((uint64_t)1 << sbits) - 1;
  const int64_t step = shares * kSliceThreads;
  bool out_of_range = false;
  for (int64_t i0 = start + c * kSliceThreads + tid; i0 < end; i0 += 8 * step) {
    uint64_t k[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) k[u] = part[min(i0 + u * step, end - 1)];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      if (i0 + u * step >= end) continue;
      if ((k[u] >> key_bits) != 0) {
        out_of_range = true;
        continue;
      }
      const uint32_t l = (uint32_t)(k[u] & lmask);
      atomicOr(&bm[l >> 5], 1u << (l & 31));
    }
  }
  if (out_of_range) atomicAdd(bad, 1u);
  __syncthreads();
  uint32_t* dst = bits + c * words + p * slice_words;
  for (int64_t e = tid; e < slice_words; e += kSliceThreads) dst[e] = bm[e];
}

// blk[b] = set bits in words [b * 4096, (b + 1) * 4096) of the bitmap, the OR
// of `shares` bitmaps laid out one after another (stored into the first when
// there are several).
__global__ void __launch_bounds__(kBmThreads)
bitmap_count_kernel(uint32_t* __restrict__ bm, int64_t words, int64_t shares,
                    uint32_t* __restrict__ blk) {
  __shared__ uint32_t ws[kBmThreads / 64 + 1];
  const int64_t base = (int64_t)blockIdx.x * kBmBlockWords + threadIdx.x;
  uint32_t s = 0;
  for (int j = 0; j < kBmBlockWords / kBmThreads; ++j) {
    const int64_t i = base + (int64_t)j * kBmThreads;
    if (i < words) {
      uint32_t v = bm[i];
      if (shares > 1) {
        for (int64_t c = 1; c < shares; ++c) v |= bm[c * words + i];
        bm[i] = v;
      }
      s += (uint32_t)__popc(v);
    }
  }
  uint32_t tot;
  block_exclusive_scan(s, ws, tot);
  if (threadIdx.x == 0) blk[blockIdx.x] = tot;
}

// Writes the key of every set bit, ascending. Each wave takes 1,024
// consecutive words, 16 per lane interleaved (word j * 64 + lane of the
// wave's run), so that the keys of one step j go to one contiguous range of
// the output: a wave scan of the lanes' popcounts places them.
__global__ void __launch_bounds__(kBmThreads)
bitmap_emit_kernel(const uint32_t* __restrict__ bm, int64_t words,
                   const uint32_t* __restrict__ blk_off, uint64_t* __restrict__ out) {
  __shared__ uint32_t wsum[kBmThreads / 64];
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int64_t base = (int64_t)blockIdx.x * kBmBlockWords + (int64_t)w * kBmWaveWords;
  uint32_t wd[kBmLaneWords];
  uint32_t t = 0;
#pragma unroll
  for (int j = 0; j < kBmLaneWords; ++j) wd[j] = bm[min(base + j * 64 + lane, words - 1)];
#pragma unroll
  for (int j = 0; j < kBmLaneWords; ++j) {
    if (base + j * 64 + lane >= words) wd[j] = 0u;
    t += (uint32_t)__popc(wd[j]);
  }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) t += __shfl_xor(t, off);
  if (lane == 0) wsum[w] = t;
  __syncthreads();
  uint32_t pos = blk_off[blockIdx.x];
  for (int v = 0; v < w; ++v) pos += wsum[v];
#pragma unroll
  for (int j = 0; j < kBmLaneWords; ++j) {
    const uint32_t c = (uint32_t)__popc(wd[j]);
    uint32_t x = c;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t y = __shfl_up(x, off);
      if (lane >= off) x += y;
    }
    uint32_t p = pos + x - c;
    const uint64_t hi = (uint64_t)(base + j * 64 + lane) << 5;
    for (uint32_t m = wd[j]; m != 0; m &= m - 1) out[p++] = hi | (uint64_t)__builtin_ctz(m);
    pos += __shfl(x, 63);
  }
}

// *out = the distinct-key count, or -1 when some key was out of range.
__global__ void store_count_kernel(const uint32_t* __restrict__ total,
                                   const uint32_t* __restrict__ bad, int64_t* __restrict__ out) {
  if (threadIdx.x == 0) *out = *bad ? -1 : (int64_t)*total;
}

__global__ void zero_word_kernel(uint32_t* __restrict__ p) {
  if (threadIdx.x == 0) *p = 0;
}

// Workspace. Radix way: two key buffers, the histograms, the digit totals,
// the tile counts. Bitmap way: the keys grouped by slice, the histograms and
// digit totals of that pass (with more than one slice), the shares' bitmaps
// and the block counts. Both end with the total and the out-of-range count.
struct IntakeWork {
  bool bitmap;
  BitmapPlan plan;
  uint64_t* a;
  uint64_t* b;
  uint32_t* hist;
  uint32_t* dtot;
  uint32_t* tile_cnt;
  uint32_t* bm;
  uint32_t* blk;
  uint32_t* total;
  uint32_t* bad;
  size_t bytes;
};

inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

IntakeWork intake_layout(int64_t n, int key_bits, void* base) {
  IntakeWork w{};
  char* p = static_cast<char*>(base);
  auto at = [p](size_t off) { return p ? p + off : nullptr; };
  const size_t tiles = (size_t)ceil_div(n, kTile);
  size_t off = 0;
  w.bitmap = bitmap_fits(n, key_bits);
  if (w.bitmap) {
    w.plan = bitmap_plan(n, key_bits);
    if (w.plan.slices > 1) {
      w.a = reinterpret_cast<uint64_t*>(at(off));
      off += align256(sizeof(uint64_t) * (size_t)n);
      w.hist = reinterpret_cast<uint32_t*>(at(off));
      off += align256(sizeof(uint32_t) * (size_t)w.plan.slices * tiles);
      w.dtot = reinterpret_cast<uint32_t*>(at(off));
      off += align256(sizeof(uint32_t) * kDigits);
    }
    w.bm = reinterpret_cast<uint32_t*>(at(off));
    off += align256(sizeof(uint32_t) * (size_t)(w.plan.words * w.plan.shares));
    w.blk = reinterpret_cast<uint32_t*>(at(off));
    off += align256(sizeof(uint32_t) * (size_t)ceil_div(w.plan.words, kBmBlockWords));
  } else {
    w.a = reinterpret_cast<uint64_t*>(at(off));
    off += align256(sizeof(uint64_t) * (size_t)n);
    w.b = reinterpret_cast<uint64_t*>(at(off));
    off += align256(sizeof(uint64_t) * (size_t)n);
    w.hist = reinterpret_cast<uint32_t*>(at(off));
    off += align256(sizeof(uint32_t) * kDigits * tiles);
    w.dtot = reinterpret_cast<uint32_t*>(at(off));
    off += align256(sizeof(uint32_t) * kDigits);
    w.tile_cnt = reinterpret_cast<uint32_t*>(at(off));
    off += align256(sizeof(uint32_t) * tiles);
  }
  w.total = reinterpret_cast<uint32_t*>(at(off));
  w.bad = reinterpret_cast<uint32_t*>(at(off + 4));
  w.bytes = off + 256;
  return w;
}

// One stable counting pass over digit (k >> shift) & (2^dbits - 1).
int radix_pass(const uint64_t* src, uint64_t* dst, int64_t n, int shift, int dbits,
               const IntakeWork& w, int key_bits, uint32_t* bad, hipStream_t s) {
  const int64_t tiles = ceil_div(n, kTile);
  const uint32_t dmask = (1u << dbits) - 1u;
  radix_hist_kernel<<<(unsigned)tiles, kSortThreads, 0, s>>>(src, n, shift, dmask, tiles, w.hist,
                                                            key_bits, bad);
  SQDK_LAUNCHED("radix_hist_kernel");
  exclusive_scan_kernel<<<dmask + 1, kScanThreads, 0, s>>>(w.hist, tiles, w.dtot);
  SQDK_LAUNCHED("exclusive_scan_kernel");
  radix_scatter_kernel<<<(unsigned)tiles, kSortThreads, 0, s>>>(src, dst, n, shift, dbits, tiles,
                                                               w.hist, w.dtot);
  SQDK_LAUNCHED("radix_scatter_kernel");
  return SQDK_OK;
}

int unique_by_bitmap(const uint64_t* keys, int64_t n, int key_bits, uint64_t* out,
                     int64_t* n_unique, const IntakeWork& w, hipStream_t s) {
  const BitmapPlan& b = w.plan;
  const int64_t blocks = ceil_div(b.words, kBmBlockWords);
  zero_word_kernel<<<1, 64, 0, s>>>(w.bad);
  SQDK_LAUNCHED("zero_word_kernel");
  const uint64_t* part = keys;
  if (b.slices > 1) {
    const int rc = radix_pass(keys, w.a, n, b.sbits, b.pbits, w, 64, nullptr, s);
    if (rc != SQDK_OK) return rc;
    part = w.a;
  }
  bitmap_slice_kernel<<<(unsigned)(b.slices * b.shares), kSliceThreads, 0, s>>>(
      part, n, key_bits, b.sbits, b.slices, b.shares, w.dtot, b.words, b.slice_words, w.bm,
      w.bad);
  SQDK_LAUNCHED("bitmap_slice_kernel");
  bitmap_count_kernel<<<(unsigned)blocks, kBmThreads, 0, s>>>(w.bm, b.words, b.shares, w.blk);
  SQDK_LAUNCHED("bitmap_count_kernel");
  exclusive_scan_kernel<<<1, kScanThreads, 0, s>>>(w.blk, blocks, w.total);
  SQDK_LAUNCHED("exclusive_scan_kernel");
  bitmap_emit_kernel<<<(unsigned)blocks, kBmThreads, 0, s>>>(w.bm, b.words, w.blk, out);
  SQDK_LAUNCHED("bitmap_emit_kernel");
  store_count_kernel<<<1, 64, 0, s>>>(w.total, w.bad, n_unique);
  SQDK_LAUNCHED("store_count_kernel");
  return SQDK_OK;
}

int unique_by_radix(const uint64_t* keys, int64_t n, int key_bits, uint64_t* out,
                    int64_t* n_unique, const IntakeWork& w, hipStream_t s) {
  const int64_t tiles = ceil_div(n, kTile);
  SQDK_REQUIRE(tiles < (1ll << 31), "too many tiles");
  const int passes = (key_bits + kDigitBits - 1) / kDigitBits;
  uint64_t* bufs[2] = {w.a, w.b};
  zero_word_kernel<<<1, 64, 0, s>>>(w.bad);
  SQDK_LAUNCHED("zero_word_kernel");
  for (int p = 0; p < passes; ++p) {
    const uint64_t* src = p == 0 ? keys : bufs[(p + 1) & 1];
    const int shift = p * kDigitBits;
    const int rc = radix_pass(src, bufs[p & 1], n, shift, std::min(kDigitBits, key_bits - shift),
                              w, key_bits, p == 0 ? w.bad : nullptr, s);
    if (rc != SQDK_OK) return rc;
  }
  const uint64_t* sorted = bufs[(passes - 1) & 1];
  unique_count_kernel<<<(unsigned)tiles, kSortThreads, 0, s>>>(sorted, n, w.tile_cnt);
  SQDK_LAUNCHED("unique_count_kernel");
  exclusive_scan_kernel<<<1, kScanThreads, 0, s>>>(w.tile_cnt, tiles, w.total);
  SQDK_LAUNCHED("exclusive_scan_kernel");
  unique_write_kernel<<<(unsigned)tiles, kSortThreads, 0, s>>>(sorted, n, w.tile_cnt, out);
  SQDK_LAUNCHED("unique_write_kernel");
  store_count_kernel<<<1, 64, 0, s>>>(w.total, w.bad, n_unique);
  SQDK_LAUNCHED("store_count_kernel");
  return SQDK_OK;
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
  pack_kernel<<<(unsigned)ceil_div(n_rows, kRowThreads), kRowThreads, 0, as_stream(stream)>>>(
      states, n_rows, num_qubits, row_stride, keys);
  SQDK_LAUNCHED("pack_kernel");
  return SQDK_OK;
}

int sqdk_intake_unique_workspace(int64_t n, int32_t key_bits, size_t* bytes) {
  SQDK_REQUIRE(bytes, "NULL output");
  SQDK_REQUIRE(n >= 1 && n < (1ll << 31), "n=%lld out of range", (long long)n);
  SQDK_REQUIRE(key_bits >= 1 && key_bits <= 64, "key_bits %d out of range", key_bits);
  *bytes = intake_layout(n, key_bits, nullptr).bytes;
  return SQDK_OK;
}

int sqdk_intake_unique(const uint64_t* keys, int64_t n, int32_t key_bits, uint64_t* out,
                       int64_t* n_unique, void* work, size_t work_bytes, void* stream) {
  SQDK_REQUIRE(keys && out && n_unique && work, "NULL pointer argument");
  SQDK_REQUIRE(n >= 1 && n < (1ll << 31), "n=%lld out of range", (long long)n);
  SQDK_REQUIRE(key_bits >= 1 && key_bits <= 64, "key_bits %d out of range", key_bits);
  const IntakeWork w = intake_layout(n, key_bits, work);
  SQDK_REQUIRE(work_bytes >= w.bytes, "intake workspace of %zu bytes below %zu", work_bytes,
               w.bytes);
  hipStream_t s = as_stream(stream);
  return w.bitmap ? unique_by_bitmap(keys, n, key_bits, out, n_unique, w, s)
                  : unique_by_radix(keys, n, key_bits, out, n_unique, w, s);
}

int sqdk_intake_unpack(const uint64_t* keys, int64_t n, int32_t num_qubits, uint8_t* states,
                       void* stream) {
  SQDK_REQUIRE(keys && states, "NULL pointer argument");
  SQDK_REQUIRE(n >= 1 && n < (1ll << 31), "n=%lld out of range", (long long)n);
  SQDK_REQUIRE(num_qubits >= 1 && num_qubits <= 64, "num_qubits %d out of range", num_qubits);
  unpack_kernel<<<(unsigned)ceil_div(n, kRowThreads), kRowThreads, 0, as_stream(stream)>>>(
      keys, n, num_qubits, states);
  SQDK_LAUNCHED("unpack_kernel");
  return SQDK_OK;
}

}  // extern "C"
