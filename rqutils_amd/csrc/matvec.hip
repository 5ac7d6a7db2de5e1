// Matrix-free Pauli-sum matvec over a sorted subspace (bucket join).
//
// The reference computes Hv = sum_g C_g o B[x_g](v) group by group: a sort
// join of [S; S ^ x_g] finds the source index of every row, a while-loop over
// the group's terms builds C_g, and a fill-mode gather applies it
// (rqutils/sqd.py:448-503, 554-587, 641-654). Here one workgroup owns a chunk
// of rows of one aligned key bucket a (keys sharing key >> m) and keeps those
// rows' keys and accumulators in registers. Off-diagonal groups are ordered by
// class, the high part h = x >> m, because every target s ^ x of a row in
// bucket a lies in bucket a ^ h. For each class the workgroup stages that one
// target bucket (its low key bits and a fine directory) in LDS with coalesced
// loads, then every group of the class costs one XOR and one short search in
// LDS per row. Hits (rare) gather v from the target bucket's slice, which is
// contiguous. Each output element is written once, by one thread, after all
// groups: no atomics, deterministic summation order.
//
// Rows are laid out wave by wave: wave w of a chunk owns its rows
// [64 R w, 64 R (w + 1)), slot r of lane l being row 64 R w + 64 r + l, so a
// short chunk (the tail of a bucket) fills whole waves and leaves the others
// idle instead of thinning out every wave.
//
// Summation: the products of one class (the groups sharing a target bucket)
// are summed into a per-row partial with FMAs; when the class ends, if the
// wave had any hit in it, every row's partial enters its running sum through
// one two-sum per part (complex64) or one add (complex128). A miss gathers 0
// (raw buffer loads return 0 past the slice) and so leaves the partial as it
// is, which keeps the hit path free of selects.
//
// Two ways to apply the hits of a group:
//  * register mode: every lane forms the group coefficient and the product
//    for all its rows, hit or not (cheap when the hits are dense enough that
//    the wave has some in most groups, and when they are so rare that the
//    wave skips most groups);
//  * compacted mode: the wave queues its hits in LDS (ballot + lane prefix
//    count), then lanes take one queued hit each, so the coefficient, gather
//    and product cost one lane per hit instead of one per row. The class
//    partials live in LDS, one per row slot. Each hit runs the same
//    arithmetic as in register mode, in the same group order, and the folds
//    happen at the same class ends, so both modes give identical bits.
#include "internal.hpp"

#include <cstring>
#include <mutex>
#include <vector>

namespace sqdk {

namespace {

constexpr int kBlock = 256;

template <typename VT> struct CoefT;
template <> struct CoefT<double2> { using type = double2; using real = double; };
template <> struct CoefT<float2> { using type = float2; using real = float; };

// Scalar launch parameters; every table is a separate __restrict__ kernel
// argument so that the compiler can read uniform entries with scalar loads.
struct MatvecParams {
  const int32_t* skip;       // whole launch skipped when *skip != 0
  const int32_t* inst_skip;  // per instance: its chunks skipped when != 0 (or NULL)
  int32_t* counter;
  int32_t n_items;
  int32_t xcd_remap;
  int32_t max_bucket;        // largest bucket over all instances
  int32_t dmax;              // largest fine_bits over all instances (LDS layout)
  int32_t split;             // workgroups sharing one row chunk (groups dealt round robin)
  int32_t threads;           // workgroup size: 256, or 512 with 4 rows per thread
  int32_t compact;           // hits applied through the per-wave LDS queue
  int64_t n;                 // rows, for the partial-sum layout
  void* partial;             // (split - 1) x n partial sums when split > 1
  sqdk_inst single;          // the instance when no descriptor array is given
};

template <typename VT>
struct MatvecTables {
  const uint64_t* keys;
  const uint32_t* bstart;
  const uint16_t* fdir;
  const uint32_t* items;
  const uint64_t* xmask;
  const int32_t* grp_off;
  const uint64_t* zmask;
  const typename CoefT<VT>::type* coef;
  const uint32_t* class_hi;
  const int32_t* class_off;
  const VT* v;
  VT* out;
  const sqdk_inst* descs;
  const uint16_t* item_inst;
  const uint32_t* fbits;     // dense lookup bitmaps (or NULL: search)
};

// Workgroup i runs on XCD i % 8; give each XCD a contiguous range of items so
// chunks of one bucket, which read the same target buckets, share an L2.
__device__ inline int remap_xcd(int i, int g) {
  constexpr int kXcd = 8;
  const int per = g / kXcd, rem = g % kXcd;
  const int xcd = i % kXcd, local = i / kXcd;
  return xcd < rem ? xcd * (per + 1) + local : rem * (per + 1) + (xcd - rem) * per + local;
}

// Byte layout of the dynamic LDS tables of one workgroup.
// The staged key slice is followed by kKeyPad entries of the key type's
// largest value, so a lifting probe up to 2^kKeyPadLog2 - 1 positions past the
// slice reads a key above every target (targets are below 2^m, m < the key
// width) and needs no clamp.
constexpr int kKeyPadLog2 = 6, kKeyPad = 1 << kKeyPadLog2;

struct LdsLayout {
  size_t pairs, occ, keys, vals, total;
};

__host__ __device__ inline LdsLayout lds_layout(int fine_bits, int max_bucket, size_t key_bytes, size_t val_bytes) {
  LdsLayout l;
  l.pairs = 0;
  l.occ = ((size_t)4 << fine_bits);
  l.keys = l.occ + 16 * sizeof(int);
  l.vals = (l.keys + key_bytes * ((size_t)max_bucket + kKeyPad) + 15) & ~(size_t)15;
  l.total = l.vals + val_bytes * (size_t)max_bucket;
  return l;
}

template <typename VT> constexpr bool kGatherFirst = sizeof(VT) == 8;

// Compacted mode: per row slot of the chunk, its class partial, one queue
// entry and its low key bits, placed after the lookup tables.
struct CompactLayout {
  size_t acc, queue, rows, total;
};

__host__ __device__ inline CompactLayout compact_layout(size_t table_bytes, int chunk_rows,
                                                        size_t acc_bytes, size_t key_bytes) {
  CompactLayout c;
  c.acc = (table_bytes + 15) & ~(size_t)15;
  c.queue = c.acc + acc_bytes * (size_t)chunk_rows;
  c.rows = c.queue + 4 * (size_t)chunk_rows;
  c.total = c.rows + key_bytes * (size_t)chunk_rows;
  return c;
}

// Orders this wave's LDS writes before its later LDS reads (other lanes read
// what a lane wrote); no cross-wave synchronisation is implied.
__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// +1 or -1 by bit 0 of a popcount: the IEEE one with bit 0 moved into the sign.
__device__ inline float parity_sign(uint32_t pc, float) {
  return __uint_as_float(0x3f800000u | (pc << 31));
}
__device__ inline double parity_sign(uint32_t pc, double) {
  return __hiloint2double((int)(0x3ff00000u | (pc << 31)), 0);
}
__device__ inline float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ inline double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }

template <typename VT> __device__ inline VT vzero();
template <> __device__ inline float2 vzero<float2>() { return make_float2(0.f, 0.f); }
template <> __device__ inline double2 vzero<double2>() { return make_double2(0.0, 0.0); }

// Gather from a target slice through a raw buffer resource covering exactly
// the slice: buffer loads return 0 for an offset at or past the resource's
// size, so a miss (index -1, a byte offset near 2^32) reads 0 without a
// clamp or a select, and its product leaves the partial sum unchanged.
constexpr int kBufferFlags = 0x00020000;  // gfx9 buffer descriptor word 3 (32-bit data)

template <typename VT>
__device__ inline __amdgpu_buffer_rsrc_t slice_rsrc(const VT* base, int count) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<VT*>(base), (short)0,
                                           count * (int)sizeof(VT), kBufferFlags);
}
__device__ inline float2 gather0(__amdgpu_buffer_rsrc_t r, int idx, float2) {
  const auto x = __builtin_amdgcn_raw_buffer_load_b64(r, (uint32_t)idx << 3, 0, 0);
  return make_float2(__uint_as_float(x[0]), __uint_as_float(x[1]));
}
__device__ inline double2 gather0(__amdgpu_buffer_rsrc_t r, int idx, double2) {
  const auto x = __builtin_amdgcn_raw_buffer_load_b128(r, (uint32_t)idx << 4, 0, 0);
  return make_double2(__hiloint2double((int)x[1], (int)x[0]),
                      __hiloint2double((int)x[3], (int)x[2]));
}

// The fields of an instance descriptor in scalar registers (the descriptor is
// the same for the whole workgroup; without this the compiler may keep it in
// vector registers and turn every loop over its ranges into a divergent one).
__device__ inline int64_t uniform_i64(int64_t x) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)x);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)((uint64_t)x >> 32));
  return (int64_t)(((uint64_t)hi << 32) | lo);
}

__device__ inline sqdk_inst uniform_inst(const sqdk_inst& d) {
  sqdk_inst u;
  u.bstart_base = uniform_i64(d.bstart_base);
  u.fdir_base = uniform_i64(d.fdir_base);
  u.class_base = __builtin_amdgcn_readfirstlane(d.class_base);
  u.n_classes = __builtin_amdgcn_readfirstlane(d.n_classes);
  u.diag_t0 = __builtin_amdgcn_readfirstlane(d.diag_t0);
  u.diag_t1 = __builtin_amdgcn_readfirstlane(d.diag_t1);
  u.bucket_bits = __builtin_amdgcn_readfirstlane(d.bucket_bits);
  u.fine_bits = __builtin_amdgcn_readfirstlane(d.fine_bits);
  u.fbits_base = uniform_i64(d.fbits_base);
  return u;
}

// sum_k c_k (-1)^{popcount(z_k & s_r)} for every row r of a thread, terms in
// index order (the order of group_coef, so results are identical: adding
// (+-1) c with one FMA rounds exactly like adding +-c). The term tables are
// read once per term for all rows.
// NARROW: every row key of the chunk is below 2^32, so the parity needs only
// the low words. Needs t1 > t0 (callers skip groups without terms).
template <typename CT, typename KT, int ROWS, bool NARROW>
__device__ inline void rows_coef_n(const uint64_t* __restrict__ zmask, const CT* __restrict__ coef,
                                   int t0, int t1, uint64_t bucket_base, const KT (&sl)[ROWS],
                                   decltype(CT::x) (&cr)[ROWS], decltype(CT::x) (&ci)[ROWS]) {
  using RT = decltype(CT::x);
  {
    // first term: (+-1) c + 0 with one FMA, as adding +-c to zero
    const uint64_t z = zmask[t0];
    const CT c = coef[t0];
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
      const uint32_t pc = NARROW
                              ? (uint32_t)__popc((uint32_t)z & ((uint32_t)bucket_base | (uint32_t)sl[r]))
                              : (uint32_t)__popcll(z & (bucket_base | (uint64_t)sl[r]));
      const RT sgn = parity_sign(pc, RT(0));
      cr[r] = fma_t(sgn, c.x, RT(0));
      ci[r] = fma_t(sgn, c.y, RT(0));
    }
  }
  for (int k = t0 + 1; k < t1; ++k) {
    const uint64_t z = zmask[k];
    const CT c = coef[k];
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
      const uint32_t pc = NARROW
                              ? (uint32_t)__popc((uint32_t)z & ((uint32_t)bucket_base | (uint32_t)sl[r]))
                              : (uint32_t)__popcll(z & (bucket_base | (uint64_t)sl[r]));
      const auto sgn = parity_sign(pc, cr[r]);
      cr[r] = fma_t(sgn, c.x, cr[r]);
      ci[r] = fma_t(sgn, c.y, ci[r]);
    }
  }
}

template <typename CT, typename KT, int ROWS>
__device__ inline void rows_coef(const uint64_t* __restrict__ zmask, const CT* __restrict__ coef,
                                 int t0, int t1, uint64_t bucket_base, bool narrow,
                                 const KT (&sl)[ROWS], decltype(CT::x) (&cr)[ROWS],
                                 decltype(CT::x) (&ci)[ROWS]) {
  if (narrow)
    rows_coef_n<CT, KT, ROWS, true>(zmask, coef, t0, t1, bucket_base, sl, cr, ci);
  else
    rows_coef_n<CT, KT, ROWS, false>(zmask, coef, t0, t1, bucket_base, sl, cr, ci);
}

// Compacted application of one group's hits for the calling wave. idx[r] is
// the hit's index in the target slice vsrc (-1: miss) for row slot
// r * 64 + lane; lq, lrow and lpart are this wave's queue, low key bits and
// class partial sums (64 * ROWS entries each). A slot is hit at most once per
// group, so the lanes of one pass never update the same partial. Returns
// whether the wave had a hit.
template <typename VT, typename KT, int ROWS>
__device__ inline bool compact_group(const int (&idx)[ROWS], int lane, uint32_t* __restrict__ lq,
                                     const KT* __restrict__ lrow, VT* __restrict__ lpart,
                                     const VT* __restrict__ vsrc,
                                     const uint64_t* __restrict__ zmask,
                                     const typename CoefT<VT>::type* __restrict__ coef, int t0,
                                     int t1, uint64_t bucket_base, bool narrow) {
  using CT = typename CoefT<VT>::type;
  using RT = typename CoefT<VT>::real;
  uint64_t mk[ROWS];
  int off[ROWS];
  int total = 0;
#pragma unroll
  for (int r = 0; r < ROWS; ++r) {
    mk[r] = __ballot(idx[r] >= 0);
    off[r] = total;
    total += __popcll(mk[r]);
  }
  if (total == 0) return false;
  // Queue position: hits of slot 0 first, then slot 1, ...; lane order within.
#pragma unroll
  for (int r = 0; r < ROWS; ++r) {
    if (idx[r] >= 0) {
      const int below = (int)__builtin_amdgcn_mbcnt_hi(
          (uint32_t)(mk[r] >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mk[r], 0u));
      lq[off[r] + below] = (uint32_t)idx[r] | ((uint32_t)(r * kWave + lane) << 16);
    }
  }
  wave_lds_sync();
  for (int j = lane; j < total; j += kWave) {
    const uint32_t e = lq[j];
    const int slot = (int)(e >> 16);
    const VT vv = vsrc[e & 0xFFFFu];
    const KT s[1] = {lrow[slot]};
    RT cr[1], ci[1];
    rows_coef<CT, KT, 1>(zmask, coef, t0, t1, bucket_base, narrow, s, cr, ci);
    VT p = lpart[slot];
    part_fma(p, cr[0], ci[0], vv);
    lpart[slot] = p;
  }
  wave_lds_sync();
  return true;
}

// End of a class with hits in the wave: every row's class partial enters its
// compensated sum (register partials, or this wave's LDS partials).
template <typename VT, int ROWS, bool COMPACT>
__device__ inline void fold_class(Acc<VT> (&acc)[ROWS], VT (&pacc)[ROWS], VT* __restrict__ lpart,
                                  int lane) {
  if constexpr (COMPACT) {
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
      VT p = lpart[r * kWave + lane];
      acc[r].fold(p);
      lpart[r * kWave + lane] = p;
    }
    wave_lds_sync();
  } else {
#pragma unroll
    for (int r = 0; r < ROWS; ++r) acc[r].fold(pacc[r]);
  }
}

// LDS address of a generic pointer into the workgroup's shared memory, and
// back: the search keeps 32-bit LDS addresses, which a read takes as is.
template <typename T>
using LdsPtr = const T __attribute__((address_space(3)))*;

template <typename T>
__device__ inline uint32_t lds_addr(const T* p) {
  return (uint32_t)(uintptr_t)(LdsPtr<T>)p;
}

template <typename T>
__device__ inline T lds_load(uint32_t addr) {
  return *(LdsPtr<T>)(uintptr_t)addr;
}

// Search of each row's target tl in the staged key slice. On entry ba holds
// the LDS address of the first key of the target's sub-bucket (staged in the
// sub-bucket table); binary lifting over 2^steps positions moves it to the
// last key <= tl: per step an add, one LDS read, a compare and a select. Keys
// past the sub-bucket are larger than the target, so ba ends on the target if
// it is present. Strides of 2^kKeyPadLog2 positions or more (a sub-bucket of
// more than 64 keys) are clamped to the slice's last key, `last`; the smaller
// ones end at most 63 positions past the slice, in the padding. hit[r]: the
// key found is the target.
template <typename KT, int ROWS>
__device__ inline void lift_search(uint32_t (&ba)[ROWS], const KT (&tl)[ROWS], int steps,
                                   uint32_t last, bool (&hit)[ROWS]) {
  int st = steps - 1;
  if (st >= kKeyPadLog2) {
#pragma unroll
    for (int r = 0; r < ROWS; ++r) ba[r] = min(ba[r], last);
    for (; st >= kKeyPadLog2; --st) {
      const uint32_t step = (uint32_t)sizeof(KT) << st;
#pragma unroll
      for (int r = 0; r < ROWS; ++r) {
        const uint32_t pa = min(ba[r] + step, last);
        ba[r] = lds_load<KT>(pa) <= tl[r] ? pa : ba[r];
      }
    }
  }
  for (; st >= 0; --st) {
    const uint32_t step = (uint32_t)sizeof(KT) << st;
    uint32_t pa[ROWS];
    KT kv[ROWS];
#pragma unroll
    for (int r = 0; r < ROWS; ++r) pa[r] = ba[r] + step;
#pragma unroll
    for (int r = 0; r < ROWS; ++r) kv[r] = lds_load<KT>(pa[r]);
#pragma unroll
    for (int r = 0; r < ROWS; ++r) ba[r] = kv[r] <= tl[r] ? pa[r] : ba[r];
  }
  KT kb[ROWS];
#pragma unroll
  for (int r = 0; r < ROWS; ++r) kb[r] = lds_load<KT>(ba[r]);
#pragma unroll
  for (int r = 0; r < ROWS; ++r) hit[r] = kb[r] == tl[r];
}

// Stages the target bucket's slice for the search: the low key bits, the
// padding, and per sub-bucket the LDS byte address of its first key; returns
// the lifting steps, ceil(log2(largest sub-bucket)). Every thread of the
// workgroup calls it (two barriers).
// Dense lookup: stages the target bucket's (occupancy bitmap, first-key
// offset) pair of every 32-value sub-bucket, kStageDense pairs per thread per
// round. No barriers: the caller places them.
constexpr int kStageDense = 4;

template <int NT>
__device__ inline void stage_dense(const uint32_t* __restrict__ fb,
                                   const uint16_t* __restrict__ fdr, uint64_t fbase, int ndir,
                                   uint2* ldir, int tid) {
  for (int e0 = tid; e0 < ndir; e0 += kStageDense * NT) {
    uint32_t bits[kStageDense], off[kStageDense];
#pragma unroll
    for (int u = 0; u < kStageDense; ++u) {
      const int e = e0 + u * NT;
      bits[u] = e < ndir ? fb[fbase + e] : 0u;
      off[u] = e < ndir ? (uint32_t)fdr[fbase + e] : 0u;
    }
#pragma unroll
    for (int u = 0; u < kStageDense; ++u) {
      const int e = e0 + u * NT;
      if (e < ndir) ldir[e] = make_uint2(bits[u], off[u]);
    }
  }
}

// The staging loops keep several loads per thread in flight: all of a round's
// loads are issued before the first LDS store, so a round costs one memory
// latency instead of one per element.
constexpr int kStageKeys = 8;  // keys per thread per round
constexpr int kStageDir = 2;   // directory entries per thread per round

template <typename KT, int NT>
__device__ inline int stage_search(const uint64_t* __restrict__ keys, uint64_t qs, int qn,
                                   uint64_t lowmask, const uint16_t* __restrict__ fdr,
                                   uint64_t fbase, int ndir, uint32_t kaddr, KT* lkeys,
                                   uint32_t* lpair, int* s_occ, int tid) {
  __syncthreads();  // the previous class is done with the LDS tables
  for (int i0 = tid; i0 < qn; i0 += kStageKeys * NT) {
    KT kv[kStageKeys];
#pragma unroll
    for (int u = 0; u < kStageKeys; ++u) {
      const int i = i0 + u * NT;
      kv[u] = i < qn ? (KT)keys[qs + i] : (KT)0;
    }
#pragma unroll
    for (int u = 0; u < kStageKeys; ++u) {
      const int i = i0 + u * NT;
      if (i < qn) lkeys[i] = (KT)(kv[u] & (KT)lowmask);
    }
  }
  for (int i = tid; i < kKeyPad; i += NT) lkeys[qn + i] = (KT)~(KT)0;
  int occ = 0;
  for (int e0 = tid; e0 < ndir; e0 += kStageDir * NT) {
    uint32_t lo[kStageDir], hi[kStageDir];
#pragma unroll
    for (int u = 0; u < kStageDir; ++u) {
      const int e = e0 + u * NT;
      lo[u] = e < ndir ? (uint32_t)fdr[fbase + e] : 0u;
      hi[u] = (e + 1 < ndir) ? (uint32_t)fdr[fbase + e + 1] : (uint32_t)qn;
    }
#pragma unroll
    for (int u = 0; u < kStageDir; ++u) {
      const int e = e0 + u * NT;
      if (e < ndir) {
        lpair[e] = kaddr + lo[u] * (uint32_t)sizeof(KT);
        occ = max(occ, (int)(hi[u] - lo[u]));
      }
    }
  }
#pragma unroll
  for (int sh = kWave / 2; sh > 0; sh >>= 1) occ = max(occ, __shfl_xor(occ, sh));
  if ((tid & (kWave - 1)) == 0) s_occ[tid / kWave] = occ;
  __syncthreads();
  int maxocc = 0;
#pragma unroll
  for (int i = 0; i < NT / kWave; ++i) maxocc = max(maxocc, s_occ[i]);
  return maxocc <= 1 ? 0 : 32 - __clz(maxocc - 1);
}

// Occupancy: at 4 rows per thread the complex64 kernel with 32-bit LDS keys
// keeps 7 waves per SIMD (at most 72 VGPRs, no spills), the others 5 (at most
// 96 VGPRs).
template <typename VT, typename KT, int ROWS>
constexpr int kMinWaves = ROWS == 4 ? (sizeof(VT) == 8 && sizeof(KT) == 4 ? 7 : 5) : 1;

template <typename VT, typename KT, int ROWS, bool STAGE_V, bool DENSE, bool COMPACT, int NT>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(kMinWaves<VT, KT, ROWS>)))
bucket_join_kernel(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ bstart,
                   const uint16_t* __restrict__ fdir, const uint32_t* __restrict__ fbits,
                   const uint32_t* __restrict__ items,
                   const uint64_t* __restrict__ xmask,
                   const int32_t* __restrict__ grp_off, const uint64_t* __restrict__ zmask,
                   const typename CoefT<VT>::type* __restrict__ coef,
                   const uint32_t* __restrict__ class_hi, const int32_t* __restrict__ class_off,
                   const VT* __restrict__ v, VT* __restrict__ out,
                   const sqdk_inst* __restrict__ descs, const uint16_t* __restrict__ item_inst,
                   MatvecParams a) {
  using CT = typename CoefT<VT>::type;
  using RT = typename CoefT<VT>::real;
  extern __shared__ __align__(16) unsigned char smem[];
  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1);
  const int wslot = (tid / kWave) * (kWave * ROWS);  // the wave's first row slot
  const int logical = a.xcd_remap ? remap_xcd(blockIdx.x, gridDim.x) : (int)blockIdx.x;
  const int w = logical / a.split;   // row chunk
  const int gs = logical % a.split;  // this workgroup's share of the groups
  if (w >= a.n_items) return;
  if (a.skip && a.skip[0] != 0) return;
  if (a.counter && blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(a.counter, 1);
  // The chunk's instance: its directories and classes (all indices are global).
  const int inst_id = item_inst ? (int)item_inst[w] : 0;
  if (a.inst_skip && a.inst_skip[inst_id] != 0) return;
  const sqdk_inst D = descs ? descs[inst_id] : a.single;
  // A mixed batch runs one launch per lookup: each skips the other's instances.
  if (descs && (D.fbits_base >= 0) != DENSE) return;
  const uint32_t* __restrict__ fb = DENSE ? fbits + D.fbits_base : fbits;
  const uint32_t* __restrict__ bst = bstart + D.bstart_base;
  const uint16_t* __restrict__ fdr = fdir + D.fdir_base;

  const int m = D.bucket_bits;
  const int d = D.fine_bits;
  const int f = m - d;
  const uint64_t lowmask = (m >= 64) ? ~0ull : ((1ull << m) - 1ull);
  const int ndir = 1 << d;
  const LdsLayout lay = lds_layout(a.dmax, a.max_bucket, sizeof(KT), STAGE_V ? sizeof(VT) : 0);
  uint32_t* lpair = reinterpret_cast<uint32_t*>(smem + lay.pairs);
  int* s_occ = reinterpret_cast<int*>(smem + lay.occ);
  KT* lkeys = reinterpret_cast<KT*>(smem + lay.keys);
  const uint32_t kaddr = lds_addr(lkeys);  // LDS address of lkeys[0]
  VT* lv = reinterpret_cast<VT*>(smem + lay.vals);
  // Dense lookup: (occupancy bitmap, first-key offset) of every 32-value
  // sub-bucket of the staged target bucket.
  uint2* ldir = reinterpret_cast<uint2*>(smem);
  // Compacted mode: this wave's class partials, hit queue and low key bits.
  const CompactLayout cl = compact_layout(DENSE ? ((size_t)8 << a.dmax) : lay.total, ROWS * NT,
                                          sizeof(VT), sizeof(KT));
  VT* lpart = reinterpret_cast<VT*>(smem + cl.acc) + wslot;
  uint32_t* lq = reinterpret_cast<uint32_t*>(smem + cl.queue) + wslot;
  KT* lrow = reinterpret_cast<KT*>(smem + cl.rows) + wslot;

  const uint32_t row0 = items[w];
  const uint64_t bucket = keys[row0] >> m;
  const uint64_t bucket_base = m >= 64 ? 0ull : (bucket << m);
  const bool narrow = ((bucket_base | lowmask) >> 32) == 0;  // all row keys below 2^32
  const uint32_t bend = bst[bucket + 1];
  const uint32_t row1 = min(bend, row0 + (uint32_t)(ROWS * NT));

  // Each row keeps only its low key bits: the high bits are the bucket's.
  KT sl[ROWS];
  bool valid[ROWS];
  Acc<VT> acc[ROWS];
  // The products of one class (the groups sharing a target bucket) are summed
  // into a partial per row, which enters the compensated sum at the class end.
  VT pacc[ROWS];
#pragma unroll
  for (int r = 0; r < ROWS; ++r) pacc[r] = vzero<VT>();
#pragma unroll
  for (int r = 0; r < ROWS; ++r) {
    const uint32_t i = row0 + wslot + r * kWave + lane;
    valid[r] = i < row1;
    sl[r] = valid[r] ? (KT)(keys[i] & lowmask) : (KT)0;
  }

  // Waves of a short chunk that hold no row skip the per-group work.
  const bool wave_active = __any(valid[0]);
  // Dense lookup: LDS address of each row's 32-value sub-bucket entry; the
  // target's entry is this XOR the group's (x >> 5) << 3.
  // Search lookup: LDS address of each row's sub-bucket entry (the target's
  // is this XOR the group's (x >> f) << 2, since (s ^ x) >> f = (s >> f) ^ (x >> f)).
  // Both tables start at the dynamic LDS base, which is address 0 (the kernel
  // declares no static LDS; launch_rows checks), so base + (off ^ y) equals
  // (base + off) ^ y and a lookup's address is one XOR.
  const uint32_t tbase = lds_addr(smem);
  uint32_t sa[ROWS];
#pragma unroll
  for (int r = 0; r < ROWS; ++r) {
    sa[r] = tbase + (DENSE ? ((uint32_t)sl[r] >> 5) << 3 : (uint32_t)(sl[r] >> f) << 2);
    // keep the precomputed address: one XOR per lookup instead of a shift and mask
    __asm__ __volatile__("" : "+v"(sa[r]));
  }

  // X = 0 group: the diagonal (first share only), one pass over its terms
  // for all rows of the thread.
  if (gs == 0 && D.diag_t1 > D.diag_t0 && wave_active) {
    RT cr[ROWS], ci[ROWS];
    rows_coef<CT, KT, ROWS>(zmask, coef, D.diag_t0, D.diag_t1, bucket_base, narrow, sl, cr, ci);
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
      if (valid[r]) acc[r].fma(cr[r], ci[r], v[row0 + wslot + r * kWave + lane]);
    }
  }
  if constexpr (COMPACT) {
    if (wave_active) {
#pragma unroll
      for (int r = 0; r < ROWS; ++r) {
        lpart[r * kWave + lane] = vzero<VT>();
        lrow[r * kWave + lane] = sl[r];
      }
      wave_lds_sync();
    }
  }

  // Groups are dealt round robin by their index within the instance, so a
  // batched launch splits every instance exactly as its own launch would.
  const int ginst = D.n_classes > 0 ? class_off[D.class_base] : 0;
  for (int c = D.class_base; c < D.class_base + D.n_classes; ++c) {
    const int g0 = class_off[c], g1 = class_off[c + 1];
    // First group of this class dealt to share gs.
    const int gfirst = g0 + ((gs - (g0 - ginst) % a.split) + a.split) % a.split;
    if (gfirst >= g1) continue;
    const uint64_t q = bucket ^ (uint64_t)class_hi[c];
    const uint32_t qs = bst[q];
    const uint32_t qe = bst[q + 1];
    if (qs == qe) continue;  // empty target bucket: no group of the class can hit
    const int qn = (int)(qe - qs);
    if constexpr (DENSE) {
      __syncthreads();  // the previous class is done with the LDS table
      stage_dense<NT>(fb, fdr, q << d, ndir, ldir, tid);
      __syncthreads();
      if (!wave_active) continue;  // the barriers above are still reached by every wave
      const VT* __restrict__ vq = v + qs;
      const __amdgpu_buffer_rsrc_t vr = slice_rsrc(vq, qn);
      bool class_hit = false;
      // The next group's X mask and term range are loaded one group ahead.
      uint64_t xm_n = xmask[gfirst];
      int gt0_n = grp_off[gfirst], gt1_n = grp_off[gfirst + 1];
      for (int g = gfirst; g < g1; g += a.split) {
        const uint64_t xm = xm_n;
        const int gt0 = gt0_n, gt1 = gt1_n;
        if (g + a.split < g1) {
          xm_n = xmask[g + a.split];
          gt0_n = grp_off[g + a.split];
          gt1_n = grp_off[g + a.split + 1];
        }
        if (gt1 <= gt0) continue;  // a group whose terms all had zero coefficients
        const uint32_t xl = (uint32_t)(xm & lowmask);
        const uint32_t xa = (xl >> 5) << 3;
        int idx[ROWS];
        bool any_hit = false;
#pragma unroll
        for (int r = 0; r < ROWS; ++r) {
          // sub-bucket of 32 values: one LDS read, a bit test and a rank
          // (bit-field extracts use the low 5 bits of t as offset and width);
          // the rank is formed for every row and masked to -1 on a miss
          const uint32_t t = (uint32_t)sl[r] ^ xl;
          const uint64_t pw = lds_load<uint64_t>(sa[r] ^ xa);
          const uint2 pr = make_uint2((uint32_t)pw, (uint32_t)(pw >> 32));
          const uint32_t rank = pr.y + (uint32_t)__popc(__builtin_amdgcn_ubfe(pr.x, 0u, t));
          const bool hit = valid[r] && __builtin_amdgcn_ubfe(pr.x, t, 1u) != 0u;
          idx[r] = (int)(rank | (hit ? 0u : 0xFFFFFFFFu));
          // the wave's decision reads idx, so both words of the entry come in one
          // 64-bit LDS read before it (the rank is not deferred to the hit path)
          any_hit |= idx[r] >= 0;
        }
        if constexpr (COMPACT) {
          class_hit |= compact_group<VT, KT, ROWS>(idx, lane, lq, lrow, lpart, vq, zmask, coef,
                                                   gt0, gt1, bucket_base, narrow);
          continue;
        }
        if (!__any(any_hit)) continue;
        class_hit = true;
        // (the dense path has registers to spare: gather first in both precisions;
        // a miss reads 0, and adding c * 0 leaves the partial unchanged)
        VT vv[ROWS];
#pragma unroll
        for (int r = 0; r < ROWS; ++r) vv[r] = gather0(vr, idx[r], vzero<VT>());
        RT cr[ROWS], ci[ROWS];
        rows_coef<CT, KT, ROWS>(zmask, coef, gt0, gt1, bucket_base, narrow, sl, cr, ci);
#pragma unroll
        for (int r = 0; r < ROWS; ++r) part_fma(pacc[r], cr[r], ci[r], vv[r]);
      }
      if (class_hit) fold_class<VT, ROWS, COMPACT>(acc, pacc, lpart, lane);
      (void)qn;
    } else {
      const int steps = stage_search<KT, NT>(keys, qs, qn, lowmask, fdr, q << d, ndir, kaddr,
                                             lkeys, lpair, s_occ, tid);
      if (STAGE_V) {
        for (int i = tid; i < qn; i += NT) lv[i] = v[qs + i];
        __syncthreads();
      }
      const uint32_t last = kaddr + (uint32_t)(qn - 1) * (uint32_t)sizeof(KT);
      if (!wave_active) continue;  // the barriers above are still reached by every wave
      const VT* __restrict__ vq = v + qs;
      const __amdgpu_buffer_rsrc_t vr = slice_rsrc(vq, qn);
      // a miss (index -1) reads 0 from either source
      auto gather = [&](int i) -> VT {
        if constexpr (STAGE_V) return i >= 0 ? lv[i] : vzero<VT>();
        return gather0(vr, i, vzero<VT>());
      };
      bool class_hit = false;
      // The next group's X mask and term range are loaded one group ahead.
      uint64_t xm_n = xmask[gfirst];
      int gt0_n = grp_off[gfirst], gt1_n = grp_off[gfirst + 1];
      for (int g = gfirst; g < g1; g += a.split) {
        const uint64_t xm = xm_n;
        const int gt0 = gt0_n, gt1 = gt1_n;
        if (g + a.split < g1) {
          xm_n = xmask[g + a.split];
          gt0_n = grp_off[g + a.split];
          gt1_n = grp_off[g + a.split + 1];
        }
        if (gt1 <= gt0) continue;  // a group whose terms all had zero coefficients
        const KT xl = (KT)(xm & lowmask);
        const uint32_t xa = (uint32_t)(xl >> f) << 2;
        KT tl[ROWS];
        uint32_t ba[ROWS];
#pragma unroll
        for (int r = 0; r < ROWS; ++r) {
          tl[r] = sl[r] ^ xl;
          ba[r] = lds_load<uint32_t>(sa[r] ^ xa);
        }
        bool hit[ROWS];
        lift_search<KT, ROWS>(ba, tl, steps, last, hit);
        // Rows past the chunk's end (sl = 0) may hit too: their products are
        // formed like the others' and never written.
        bool any_hit = false;
#pragma unroll
        for (int r = 0; r < ROWS; ++r) any_hit |= hit[r];
        int idx[ROWS];
#pragma unroll
        for (int r = 0; r < ROWS; ++r)
          idx[r] = hit[r] ? (int)((ba[r] - kaddr) / (uint32_t)sizeof(KT)) : -1;
        if constexpr (COMPACT) {
          class_hit |= compact_group<VT, KT, ROWS>(idx, lane, lq, lrow, lpart, STAGE_V ? lv : vq,
                                                   zmask, coef, gt0, gt1, bucket_base, narrow);
          continue;
        }
        if (!__any(any_hit)) continue;
        class_hit = true;
        // Hits: gather (a miss reads 0, so its product leaves the partial
        // unchanged) and one pass over the group's terms for all rows.
        // complex64 gathers first to hide the load behind the terms; complex128
        // gathers after them, which keeps 16 fewer VGPRs live (occupancy).
        VT vv[ROWS];
        if constexpr (kGatherFirst<VT>) {
#pragma unroll
          for (int r = 0; r < ROWS; ++r) vv[r] = gather(idx[r]);
        }
        RT cr[ROWS], ci[ROWS];
        rows_coef<CT, KT, ROWS>(zmask, coef, gt0, gt1, bucket_base, narrow, sl, cr, ci);
#pragma unroll
        for (int r = 0; r < ROWS; ++r) {
          if constexpr (!kGatherFirst<VT>) vv[r] = gather(idx[r]);
          part_fma(pacc[r], cr[r], ci[r], vv[r]);
        }
      }
      if (class_hit) fold_class<VT, ROWS, COMPACT>(acc, pacc, lpart, lane);
    }
  }

  // Shares other than the first write partial sums for the reduction pass.
  VT* __restrict__ dst = gs == 0 ? out : reinterpret_cast<VT*>(a.partial) + (int64_t)(gs - 1) * a.n;
#pragma unroll
  for (int r = 0; r < ROWS; ++r) {
    if (valid[r]) dst[row0 + wslot + r * kWave + lane] = acc[r].result();
  }
}

// out[i] += sum of the partial sums of shares 1..split-1, in share order.
template <typename VT>
__global__ void __launch_bounds__(kBlock)
split_reduce_kernel(int64_t n, int nparts, const VT* __restrict__ partial, VT* __restrict__ out,
                    const int32_t* __restrict__ skip) {
  if (skip && skip[0] != 0) return;
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock) {
    Acc<VT> acc;
    const VT o = out[i];
    acc.fma(1, 0, o);
    for (int k = 0; k < nparts; ++k) acc.fma(1, 0, partial[(int64_t)k * n + i]);
    out[i] = acc.result();
  }
}

// The lookups fold the LDS base of the tables into each row's address, which
// needs the dynamic LDS to start at address 0: the kernel must declare no
// static LDS. Checked once per kernel.
inline int check_lds_base(const void* kern, const char* name) {
  static std::mutex mu;
  static std::vector<const void*> ok;
  std::lock_guard<std::mutex> lock(mu);
  for (const void* k : ok)
    if (k == kern) return SQDK_OK;
  hipFuncAttributes attr{};
  SQDK_HIP(hipFuncGetAttributes(&attr, kern));
  if (attr.sharedSizeBytes != 0)
    return fail(SQDK_EHIP, "%s declares %zu bytes of static LDS; its lookups need none", name,
                attr.sharedSizeBytes);
  ok.push_back(kern);
  return SQDK_OK;
}

template <typename VT, typename KT, int ROWS, bool SV, bool DENSE, bool COMPACT, int NT = kBlock>
int launch_rows(const MatvecTables<VT>& t, const MatvecParams& args, size_t lds, hipStream_t s) {
  // (t.descs / t.item_inst are NULL for a single instance)
  auto kern = bucket_join_kernel<VT, KT, ROWS, SV, DENSE, COMPACT, NT>;
  SQDK_TRY(check_lds_base(reinterpret_cast<const void*>(kern), "bucket_join_kernel"));
  if (lds > 65536) {
    SQDK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  }
  const int64_t grid = (int64_t)args.n_items * args.split;
  if (grid >= (1ll << 31)) return fail(SQDK_ERANGE, "grid of %lld workgroups too large", (long long)grid);
  kern<<<(unsigned)grid, NT, lds, s>>>(t.keys, t.bstart, t.fdir, t.fbits, t.items, t.xmask,
                                           t.grp_off,
                                           t.zmask, t.coef, t.class_hi, t.class_off, t.v, t.out,
                                           t.descs, t.item_inst, args);
  SQDK_LAUNCHED("bucket_join_kernel");
  return SQDK_OK;
}

// After every launch of a matvec: out += the other shares' partial sums.
template <typename VT>
int reduce_split(const MatvecTables<VT>& t, const MatvecParams& args, hipStream_t s) {
  if (args.split <= 1) return SQDK_OK;
  const unsigned g = (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(args.n, kBlock), 4096));
  split_reduce_kernel<VT><<<g, kBlock, 0, s>>>(args.n, args.split - 1,
                                               reinterpret_cast<const VT*>(args.partial), t.out,
                                               args.skip);
  SQDK_LAUNCHED("split_reduce_kernel");
  return SQDK_OK;
}

template <typename VT, typename KT, bool SV, bool DENSE, bool COMPACT>
int launch_mode(const MatvecTables<VT>& t, const MatvecParams& args, int rows, size_t lds,
                hipStream_t s) {
  // 512-thread workgroups (2,048-row chunks) exist for 4 rows per thread only.
  if (args.threads == 512) {
    if (rows != 4) return fail(SQDK_EINVAL, "512-thread chunks need 4 rows per thread");
    return launch_rows<VT, KT, 4, SV, DENSE, COMPACT, 512>(t, args, lds, s);
  }
  switch (rows) {
    case 4: return launch_rows<VT, KT, 4, SV, DENSE, COMPACT>(t, args, lds, s);
    case 8: return launch_rows<VT, KT, 8, SV, DENSE, COMPACT>(t, args, lds, s);
    case 16:
      if constexpr (!COMPACT) return launch_rows<VT, KT, 16, SV, DENSE, false>(t, args, lds, s);
      return fail(SQDK_EINVAL, "compacted hits need 4 or 8 rows per thread");
    default: return fail(SQDK_EINVAL, "rows_per_thread %d unsupported (4, 8 or 16)", rows);
  }
}

// LDS bytes of one workgroup: the lookup tables, plus the per-row areas of the
// compacted mode.
template <typename VT>
size_t lds_bytes(size_t table_bytes, const MatvecParams& args, int rows, int key_bytes) {
  if (!args.compact) return table_bytes;
  return compact_layout(table_bytes, rows * args.threads, sizeof(VT), key_bytes).total;
}

template <typename VT, typename KT, bool SV, bool DENSE = false>
int launch_kt(const MatvecTables<VT>& t, const MatvecParams& args, int rows, size_t table_bytes,
              hipStream_t s) {
  const size_t lds = lds_bytes<VT>(table_bytes, args, rows, (int)sizeof(KT));
  if (lds > 160 * 1024)
    return fail(SQDK_ERANGE, "LDS use of %zu bytes exceeds 160 KiB (tables %zu bytes%s)", lds,
                table_bytes, args.compact ? ", compacted hits" : "");
  if (args.compact) return launch_mode<VT, KT, SV, DENSE, true>(t, args, rows, lds, s);
  return launch_mode<VT, KT, SV, DENSE, false>(t, args, rows, lds, s);
}

template <typename VT>
int launch_common(const MatvecTables<VT>& t, MatvecParams& args, int narrow_keys, int rows,
                  int stage_v, hipStream_t s) {
  if (t.fbits) {
    // Dense lookup: only the (bitmap, offset) table of the target bucket is staged.
    if (stage_v) return fail(SQDK_EINVAL, "v staging is not available with the dense lookup");
    return launch_kt<VT, uint32_t, false, true>(t, args, rows, (size_t)8 << args.dmax, s);
  }
  const LdsLayout lay = lds_layout(args.dmax, args.max_bucket, narrow_keys ? 4 : 8,
                                   stage_v ? sizeof(VT) : 0);
  if (stage_v)
    return narrow_keys ? launch_kt<VT, uint32_t, true>(t, args, rows, lay.total, s)
                       : launch_kt<VT, uint64_t, true>(t, args, rows, lay.total, s);
  return narrow_keys ? launch_kt<VT, uint32_t, false>(t, args, rows, lay.total, s)
                     : launch_kt<VT, uint64_t, false>(t, args, rows, lay.total, s);
}

template <typename VT>
int launch_vt(const sqdk_ham* h, const sqdk_space* sp, const void* coef, const void* v,
              void* out, int rows, int xcd, int stage_v, int split, int compact,
              void* work, hipStream_t s, const int32_t* skip, int32_t* counter) {
  MatvecTables<VT> t;
  t.keys = sp->keys;
  t.bstart = sp->bstart;
  t.fdir = sp->fdir;
  t.items = sp->items;
  t.xmask = h->xmask;
  t.grp_off = h->grp_off;
  t.zmask = h->zmask;
  t.coef = reinterpret_cast<const typename CoefT<VT>::type*>(coef);
  t.class_hi = h->class_hi;
  t.class_off = h->class_off;
  t.v = reinterpret_cast<const VT*>(v);
  t.out = reinterpret_cast<VT*>(out);
  t.descs = nullptr;
  t.item_inst = nullptr;
  t.fbits = sp->fbits;
  if (sp->fbits && (sp->fine_bits != sp->bucket_bits - 5 || sp->bucket_bits > 20))
    return fail(SQDK_EINVAL, "dense lookup needs fine_bits = bucket_bits - 5 <= 15 (got %d, %d)",
                sp->fine_bits, sp->bucket_bits);
  MatvecParams args;
  args.skip = skip;
  args.inst_skip = nullptr;
  args.counter = counter;
  args.n_items = sp->n_items;
  args.xcd_remap = xcd;
  args.max_bucket = sp->max_bucket;
  args.dmax = sp->fine_bits;
  args.split = split;
  args.threads = sp->chunk_rows / rows;
  args.compact = compact;
  args.n = sp->n;
  args.partial = work;
  args.single.bstart_base = 0;
  args.single.fdir_base = 0;
  args.single.class_base = 0;
  args.single.n_classes = h->n_classes;
  args.single.diag_t0 = h->diag_t0;
  args.single.diag_t1 = h->diag_t1;
  args.single.bucket_bits = sp->bucket_bits;
  args.single.fine_bits = sp->fine_bits;
  args.single.fbits_base = 0;
  SQDK_TRY(launch_common<VT>(t, args, sp->bucket_bits < 32, rows, stage_v, s));
  return reduce_split<VT>(t, args, s);
}

template <typename VT>
int launch_batch(const sqdk_batch* b, const void* coef, const void* v, void* out, int rows,
                 int xcd, int stage_v, int split, int compact, void* work, hipStream_t s,
                 const int32_t* skip, const int32_t* inst_skip, int32_t* counter) {
  MatvecTables<VT> t;
  t.keys = b->keys;
  t.bstart = b->bstart;
  t.fdir = b->fdir;
  t.items = b->items;
  t.xmask = b->xmask;
  t.grp_off = b->grp_off;
  t.zmask = b->zmask;
  t.coef = reinterpret_cast<const typename CoefT<VT>::type*>(coef);
  t.class_hi = b->class_hi;
  t.class_off = b->class_off;
  t.v = reinterpret_cast<const VT*>(v);
  t.out = reinterpret_cast<VT*>(out);
  t.descs = b->inst;
  t.item_inst = b->item_inst;
  t.fbits = nullptr;
  MatvecParams args;
  args.skip = skip;
  args.inst_skip = inst_skip;
  args.counter = counter;
  args.n_items = b->n_items;
  args.xcd_remap = xcd;
  args.max_bucket = b->max_bucket;
  args.dmax = b->max_fine_bits;
  args.split = split;
  args.threads = b->chunk_rows / rows;
  args.compact = compact;
  args.n = b->n;
  args.partial = work;
  memset(&args.single, 0, sizeof(args.single));
  if (b->fbits) {
    // dense-lookup instances first (32-bit LDS keys), then the searched ones
    MatvecTables<VT> td = t;
    td.fbits = b->fbits;
    MatvecParams ad = args;
    ad.dmax = b->dense_fine_bits;
    ad.counter = nullptr;
    SQDK_TRY(launch_common<VT>(td, ad, 1, rows, 0, s));
  }
  SQDK_TRY(launch_common<VT>(t, args, b->max_bucket_bits < 32, rows, stage_v, s));
  return reduce_split<VT>(t, args, s);
}

// ---- cached hit table ---------------------------------------------------------
//
// The hits of H over the subspace (row i, source row j, coefficient C_g(s_i)
// of the group g with s_j = s_i ^ x_g), found once by the same bucket join,
// then applied by a sliced sparse product. Each window of 256 consecutive
// rows of a join chunk is sorted by hit count, most first (rows of equal
// count keep their order, and slots past the chunk's rows come last), and
// every 64 consecutive sorted slots form a slice: slot_row[64 s + l] is the
// row of lane l of slice s. A slice's entries are column-major, entry k of
// lane l at base + 64 k + l, for k below the slice's width (its first row's
// count); shorter rows are padded with column kHitPad and coefficient 0, and
// the product reads only a row's own entries (row_len). The sort keeps the
// padding small where counts vary (on the 28-qubit chain 1.39x the hits, 2.16x
// in the join's row order) and out of the cache lines of live entries, and
// puts rows without hits into zero-width slices, still in row order. Hits of
// a row are stored in class and group order. The X = 0 group is kept as a
// separate diagonal vector.
//
// Build workspace (optional). The counting pass also records
//  * per chunk and class, which of the workgroup's four waves had a hit (4
//    bits per class, 8 classes per 32-bit word, cstride words per chunk),
//    followed by one word whose bit w says that wave w has a row with more
//    than kShortHits hits;
//  * per row, the source rows of its first kShortHits hits.
// The fill pass then takes the entries of a wave whose rows all have at most
// kShortHits hits from that record (the group of a hit follows from the X mask
// s ^ t, the groups being sorted by X mask) and searches only in the other
// waves, and there only the classes they hit. The lookups are deterministic,
// so the skipped work would have found nothing and the entries are the same.
// Where hits are rare (0.1 per row on the 40-qubit sum: about 1 in 10^4 rows
// with more than two, 20% of the (chunk, class) pairs with any) this removes
// nearly all of the fill pass's join.
//
// (A layout with one step per group and a lane mask, which keeps the join's
// gather locality, measured slower: 7.6 hits per step on the 28-qubit chain,
// 1.1 on the 40-qubit sum, and each step costs a wave iteration.)

constexpr uint32_t kHitPad = 0xFFFFFFFFu;

// Rows of one sort window: 256 consecutive rows of a chunk are sorted by hit
// count among themselves. Measured (MI355X, hit product, 64-row slices):
// windows of 64 / 128 / 256 / 1024 rows give 1.49 / 1.38 / 1.33 / 1.85 ms on
// the 28-qubit chain (complex64; 1.76 ms unsorted) and 0.141 / 0.137 / 0.140 /
// 0.145 ms on the 40-qubit sum (0.150 unsorted). Within a slice the sort makes
// the lanes still walking their hits contiguous, so the padding of shorter
// rows stops sharing cache lines with live entries; wider windows also cut
// the padding itself, until rows from far apart in the chunk share a slice
// and the gathers lose their locality.
constexpr int kHitSortLog2 = 8;

// Hits per row that the counting pass records for the fill (build workspace).
constexpr int kShortHits = 2;

// MODE 0: every row's hit count into rlen, then the chunk's sort: the row of
// each slot into slot_row and each slice's entry count (64 x width) into
// sbase[sid]. MODE 1 (sbase now holds each slice's first entry): the entries
// at the slots slot_row gives, the padding and the diagonal.
// Occupancy of the table's passes at 4 rows per thread with 32-bit LDS keys: 7
// waves per SIMD (at most 72 VGPRs; the LDS tables allow 7 workgroups per CU),
// as in the join, except the complex128 fill (6: it needs up to 80 VGPRs).
// Measured on the 40-qubit sum: counting pass 25.3 -> 23.1 ms against no hint.
template <typename VT, typename KT, int ROWS, int MODE>
constexpr int kTableWaves =
    (ROWS == 4 && sizeof(KT) == 4) ? ((MODE == 1 && sizeof(VT) == 16) ? 6 : 7) : 1;

template <typename VT, typename KT, int ROWS, bool DENSE, int MODE>
__global__ void __launch_bounds__(kBlock)
__attribute__((amdgpu_waves_per_eu(kTableWaves<VT, KT, ROWS, MODE>)))
hit_table_kernel(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ bstart,
                 const uint16_t* __restrict__ fdir, const uint32_t* __restrict__ fbits,
                 const uint32_t* __restrict__ items, const uint64_t* __restrict__ xmask,
                 const int32_t* __restrict__ grp_off, const uint64_t* __restrict__ zmask,
                 const typename CoefT<VT>::type* __restrict__ coef,
                 const uint32_t* __restrict__ class_hi, const int32_t* __restrict__ class_off,
                 const sqdk_inst* __restrict__ descs, const uint16_t* __restrict__ item_inst,
                 sqdk_inst single, int dmax, int max_bucket, int sort_log2,
                 uint32_t* __restrict__ slot_row, uint64_t* __restrict__ sbase,
                 uint32_t* __restrict__ col, typename CoefT<VT>::type* __restrict__ hcoef,
                 VT* __restrict__ diag, uint32_t* __restrict__ rlen,
                 uint32_t* __restrict__ cmask, int cstride, uint32_t* __restrict__ rowhits) {
  using CT = typename CoefT<VT>::type;
  using RT = typename CoefT<VT>::real;
  extern __shared__ __align__(16) unsigned char smem[];
  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1);
  // (the wave index held in a scalar register: branches on it stay uniform)
  const int wave = __builtin_amdgcn_readfirstlane(tid / kWave);
  const int wslot = wave * (kWave * ROWS);
  // chunk: consecutive chunks (of one bucket, reading the same target buckets)
  // on one XCD, whose L2 they then share, as in the join
  const int w = remap_xcd(blockIdx.x, gridDim.x);
  // The chunk's instance (a batch) or the single one; a mixed batch runs one
  // launch per lookup, each leaving the other's chunks alone.
  const sqdk_inst D = uniform_inst(descs ? descs[item_inst[w]] : single);
  if (descs && (D.fbits_base >= 0) != DENSE) return;
  const uint32_t* __restrict__ bstart_i = bstart + D.bstart_base;
  const uint16_t* __restrict__ fdir_i = fdir + D.fdir_base;
  const uint32_t* __restrict__ fbits_i = DENSE ? fbits + D.fbits_base : fbits;
  const int m = D.bucket_bits;
  const int d = D.fine_bits;
  const int f = m - d;
  const uint64_t lowmask = (m >= 64) ? ~0ull : ((1ull << m) - 1ull);
  const int ndir = 1 << d;
  const LdsLayout lay = lds_layout(dmax, max_bucket, sizeof(KT), 0);
  uint32_t* lpair = reinterpret_cast<uint32_t*>(smem + lay.pairs);
  int* s_occ = reinterpret_cast<int*>(smem + lay.occ);
  KT* lkeys = reinterpret_cast<KT*>(smem + lay.keys);
  const uint32_t kaddr = lds_addr(lkeys);  // LDS address of lkeys[0]
  uint2* ldir = reinterpret_cast<uint2*>(smem);

  const uint32_t row0 = items[w];
  const uint64_t bucket = keys[row0] >> m;
  const uint64_t bucket_base = m >= 64 ? 0ull : (bucket << m);
  const bool narrow = ((bucket_base | lowmask) >> 32) == 0;
  const uint32_t row1 = min(bstart_i[bucket + 1], row0 + (uint32_t)(ROWS * kBlock));
  uint32_t sa[ROWS];
  // (the sub-bucket table starts at the dynamic LDS base, address 0: see the
  // join kernel)
  const uint32_t tbase = lds_addr(smem);
  KT sl[ROWS];
  bool valid[ROWS];
  uint32_t cnt[ROWS];
  uint64_t ebase[ROWS];
  constexpr int P = ROWS * kBlock;                    // slots of the chunk
  const int64_t slot0 = (int64_t)w * P;               // its first slot (64 per slice)
  // this chunk's class-mask words and wave flags (counting pass: cleared here;
  // the class loop's first barrier orders the clearing before any atomicOr)
  uint32_t* __restrict__ cm = cmask ? cmask + (int64_t)w * (cstride + 1) : nullptr;
  const int ncm = cmask ? 8 * cstride : 0;            // classes the mask covers
  if constexpr (MODE == 0) {
    if (cm)
      for (int j = tid; j <= cstride; j += kBlock) cm[j] = 0u;
  }
  // counting pass: the source rows of each row's first kShortHits hits
  uint32_t hsrc[ROWS][kShortHits];
#pragma unroll
  for (int r = 0; r < ROWS; ++r)
#pragma unroll
    for (int j = 0; j < kShortHits; ++j) hsrc[r][j] = 0u;
  if constexpr (MODE == 1) {
    // the sorted slot of each of the chunk's rows (LDS is free before the classes)
    uint16_t* inv = reinterpret_cast<uint16_t*>(smem);
    for (int q = tid; q < P; q += kBlock) {
      const uint32_t rr = slot_row[slot0 + q];
      if (rr != kHitPad) inv[rr - row0] = (uint16_t)q;
    }
    __syncthreads();
  }
#pragma unroll
  for (int r = 0; r < ROWS; ++r) {
    const uint32_t i = row0 + wslot + r * kWave + lane;
    valid[r] = i < row1;
    sl[r] = valid[r] ? (KT)(keys[i] & lowmask) : (KT)0;
    sa[r] = DENSE ? 0u : tbase + ((uint32_t)(sl[r] >> f) << 2);  // the row's sub-bucket entry
    cnt[r] = 0;
    ebase[r] = 0;
    if constexpr (MODE == 1) {
      if (valid[r]) {
        const int q = reinterpret_cast<const uint16_t*>(smem)[wslot + r * kWave + lane];
        ebase[r] = sbase[(slot0 + q) / kWave] + (uint64_t)(q % kWave);
      }
    }
  }
  const bool wave_active = __any(valid[0]);
  if (MODE == 1 && wave_active) {
    RT cr[ROWS], ci[ROWS];
#pragma unroll
    for (int r = 0; r < ROWS; ++r) cr[r] = ci[r] = RT(0);
    if (D.diag_t1 > D.diag_t0)
      rows_coef<CT, KT, ROWS>(zmask, coef, D.diag_t0, D.diag_t1, bucket_base, narrow, sl, cr, ci);
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
      VT dv;
      dv.x = cr[r];
      dv.y = ci[r];
      if (valid[r]) diag[row0 + wslot + r * kWave + lane] = dv;
    }
  }
  // (class range and class tables read into scalar registers: the loop and its
  // branches stay uniform)
  const int cbase = __builtin_amdgcn_readfirstlane(D.class_base);
  const int cend = cbase + __builtin_amdgcn_readfirstlane(D.n_classes);
  // fill pass: waves whose rows all have at most kShortHits hits (flag bit
  // clear) take their entries from the counting pass's record
  const uint32_t ovf = (MODE == 1 && cm) ? cm[cstride] : 0xFu;
  if (MODE == 1 && wave_active && ((ovf >> wave) & 1u) == 0u) {
    // the instance's off-diagonal groups, sorted by X mask
    const int gb = class_off[cbase], ge = class_off[cend];
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
      if (!valid[r]) continue;
      const uint32_t row = row0 + wslot + r * kWave + lane;
      const uint32_t nh = rlen[row];
      const uint64_t s_key = bucket_base | (uint64_t)sl[r];
#pragma unroll
      for (int j = 0; j < kShortHits; ++j) {
        if ((uint32_t)j >= nh) break;
        const uint32_t src = rowhits[(uint64_t)kShortHits * row + j];
        const uint64_t x = s_key ^ keys[src];
        int lo = gb, hi = ge;  // first group with X mask >= x: the hit's group
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          if (xmask[mid] < x) lo = mid + 1;
          else hi = mid;
        }
        RT cr1[1], ci1[1];
        const KT s1[1] = {sl[r]};
        rows_coef<CT, KT, 1>(zmask, coef, grp_off[lo], grp_off[lo + 1], bucket_base, narrow, s1,
                             cr1, ci1);
        const uint64_t e = ebase[r] + (uint64_t)j * kWave;
        col[e] = src;
        CT c;
        c.x = cr1[0];
        c.y = ci1[0];
        hcoef[e] = c;
      }
      cnt[r] = nh;
    }
  }
  // record the hits of one group (idx[r] = index in the target slice or -1)
  auto record = [&](const int (&idx)[ROWS], uint32_t qs, int gt0, int gt1) {
    if constexpr (MODE == 0) {
#pragma unroll
      for (int r = 0; r < ROWS; ++r) {
        const uint32_t src = qs + (uint32_t)idx[r];
#pragma unroll
        for (int j = 0; j < kShortHits; ++j)
          hsrc[r][j] = (idx[r] >= 0 && cnt[r] == (uint32_t)j) ? src : hsrc[r][j];
        cnt[r] += idx[r] >= 0 ? 1u : 0u;
      }
    } else {
      RT cr[ROWS], ci[ROWS];
      rows_coef<CT, KT, ROWS>(zmask, coef, gt0, gt1, bucket_base, narrow, sl, cr, ci);
#pragma unroll
      for (int r = 0; r < ROWS; ++r) {
        if (idx[r] >= 0) {
          // ebase already holds the slot's lane within its slice
          const uint64_t e = ebase[r] + (uint64_t)cnt[r] * kWave;
          col[e] = qs + (uint32_t)idx[r];
          CT c;
          c.x = cr[r];
          c.y = ci[r];
          hcoef[e] = c;
          ++cnt[r];
        }
      }
    }
  };
  if constexpr (MODE == 0) {
    if (cm) __syncthreads();  // the mask words are clear before any class records a hit
  }
  // fill pass: the chunk's first 64 mask words sit one per lane, so a class's
  // word is a register read (the rest, if any, are loaded when reached)
  uint32_t mreg = 0xFFFFFFFFu;
  if constexpr (MODE == 1) {
    if (cm && lane < cstride) mreg = cm[lane];
  }
  for (int c = cbase; c < cend; ++c) {
    const int lc = c - cbase;  // the class within the instance
    // fill pass: the waves of this chunk that had hits in this class (uniform
    // over the workgroup; all four where the mask does not cover the class),
    // tested before the class's tables are read
    uint32_t cwaves = 0xFu;
    if constexpr (MODE == 1) {
      if (lc < ncm) {
        const int wi = lc >> 3;
        const uint32_t word = wi < kWave ? (uint32_t)__builtin_amdgcn_readlane((int)mreg, wi)
                                         : cm[wi];
        cwaves = (word >> (4 * (lc & 7))) & 0xFu;
      }
      cwaves &= ovf;               // only waves with a long row search
      if (cwaves == 0u) continue;  // no such wave of the chunk has a hit here
    }
    const int g0 = __builtin_amdgcn_readfirstlane(class_off[c]);
    const int g1 = __builtin_amdgcn_readfirstlane(class_off[c + 1]);
    if (g0 >= g1) continue;
    const uint64_t q = bucket ^ (uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane(
                                    (int)class_hi[c]);
    const uint32_t qs = (uint32_t)__builtin_amdgcn_readfirstlane((int)bstart_i[q]);
    const uint32_t qe = (uint32_t)__builtin_amdgcn_readfirstlane((int)bstart_i[q + 1]);
    if (qs == qe) continue;
    const int qn = (int)(qe - qs);
    // counting pass: whether this wave had a hit in the class
    bool class_any = false;
    if constexpr (DENSE) {
      __syncthreads();
      stage_dense<kBlock>(fbits_i, fdir_i, q << d, ndir, ldir, tid);
      __syncthreads();
      if (!wave_active || ((cwaves >> wave) & 1u) == 0u) continue;
      // the next group's X mask and term range are loaded one group ahead
      uint64_t xm_n = xmask[g0];
      int gt0_n = grp_off[g0], gt1_n = grp_off[g0 + 1];
      for (int g = g0; g < g1; ++g) {
        const uint64_t xm = xm_n;
        const int gt0 = gt0_n, gt1 = gt1_n;
        if (g + 1 < g1) {
          xm_n = xmask[g + 1];
          gt0_n = gt1;
          gt1_n = grp_off[g + 2];
        }
        if (gt1 <= gt0) continue;
        const uint32_t xl = (uint32_t)(xm & lowmask);
        int idx[ROWS];
        bool any_hit = false;
#pragma unroll
        for (int r = 0; r < ROWS; ++r) {
          const uint32_t t = (uint32_t)sl[r] ^ xl;
          const uint2 pr = ldir[t >> 5];
          const uint32_t rank = pr.y + (uint32_t)__popc(__builtin_amdgcn_ubfe(pr.x, 0u, t));
          const bool hit = valid[r] && __builtin_amdgcn_ubfe(pr.x, t, 1u) != 0u;
          idx[r] = hit ? (int)rank : -1;
          any_hit |= hit;
        }
        if (__any(any_hit)) {
          record(idx, qs, gt0, gt1);
          class_any = true;
        }
      }
    } else {
      const int steps = stage_search<KT, kBlock>(keys, qs, qn, lowmask, fdir_i, q << d, ndir,
                                                 kaddr, lkeys, lpair, s_occ, tid);
      const uint32_t last = kaddr + (uint32_t)(qn - 1) * (uint32_t)sizeof(KT);
      if (!wave_active || ((cwaves >> wave) & 1u) == 0u) continue;
      uint64_t xm_n = xmask[g0];
      int gt0_n = grp_off[g0], gt1_n = grp_off[g0 + 1];
      for (int g = g0; g < g1; ++g) {
        const uint64_t xm = xm_n;
        const int gt0 = gt0_n, gt1 = gt1_n;
        if (g + 1 < g1) {
          xm_n = xmask[g + 1];
          gt0_n = gt1;
          gt1_n = grp_off[g + 2];
        }
        if (gt1 <= gt0) continue;
        const KT xl = (KT)(xm & lowmask);
        const uint32_t xa = (uint32_t)(xl >> f) << 2;
        KT tl[ROWS];
        uint32_t ba[ROWS];
#pragma unroll
        for (int r = 0; r < ROWS; ++r) {
          tl[r] = sl[r] ^ xl;
          ba[r] = lds_load<uint32_t>(sa[r] ^ xa);
        }
        bool hit[ROWS];
        lift_search<KT, ROWS>(ba, tl, steps, last, hit);
        // only the chunk's rows are recorded; the indices are formed in the
        // rare case of a hit in the wave (the test itself is on lane masks)
        bool any_hit = false;
#pragma unroll
        for (int r = 0; r < ROWS; ++r) {
          hit[r] = hit[r] && valid[r];
          any_hit |= hit[r];
        }
        if (__any(any_hit)) {
          int idx[ROWS];
#pragma unroll
          for (int r = 0; r < ROWS; ++r)
            idx[r] = hit[r] ? (int)((ba[r] - kaddr) / (uint32_t)sizeof(KT)) : -1;
          record(idx, qs, gt0, gt1);
          class_any = true;
        }
      }
    }
    if constexpr (MODE == 0) {
      if (class_any && lane == 0 && lc < ncm)
        atomicOr(&cm[lc >> 3], (1u << wave) << (4 * (lc & 7)));
    }
  }
  if constexpr (MODE == 0) {
    // Every row's count, then a bitonic sort of the chunk's slots in LDS by
    // key = window << 44 | (2^32 - 1 - count) << 12 | local row, window =
    // local row >> sort_log2 (ascending: each window of 2^sort_log2 rows on
    // its own, most hits first, equal counts in row order; slots without a
    // row carry count 0 and come after every row of their window since their
    // local index is larger).
#pragma unroll
    for (int r = 0; r < ROWS; ++r)
      if (valid[r]) rlen[row0 + wslot + r * kWave + lane] = cnt[r];
    if (cm) {
      bool longrow = false;
#pragma unroll
      for (int r = 0; r < ROWS; ++r) {
        const uint32_t row = row0 + wslot + r * kWave + lane;
        if (valid[r]) {
#pragma unroll
          for (int j = 0; j < kShortHits; ++j)
            rowhits[(uint64_t)kShortHits * row + j] = hsrc[r][j];
        }
        longrow |= valid[r] && cnt[r] > (uint32_t)kShortHits;
      }
      if (__any(longrow) && lane == 0) atomicOr(&cm[cstride], 1u << wave);
    }
    __syncthreads();  // the class loop is done with the lookup tables
    uint64_t* sk = reinterpret_cast<uint64_t*>(smem);
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
      const int loc = wslot + r * kWave + lane;
      const uint64_t c = valid[r] ? cnt[r] : 0u;
      sk[loc] = ((uint64_t)(loc >> sort_log2) << 44) | ((0xFFFFFFFFull - c) << 12) |
                (uint64_t)loc;
    }
    __syncthreads();
    for (int size = 2; size <= P; size <<= 1) {
      for (int stride = size >> 1; stride > 0; stride >>= 1) {
        for (int t = tid; t < P / 2; t += kBlock) {
          const int lo = 2 * t - (t & (stride - 1));
          const int hi = lo + stride;
          const bool up = (lo & size) == 0;
          const uint64_t a0 = sk[lo], a1 = sk[hi];
          if ((a0 > a1) == up) {
            sk[lo] = a1;
            sk[hi] = a0;
          }
        }
        __syncthreads();
      }
    }
    const int nrow = (int)(row1 - row0);
    for (int q = tid; q < P; q += kBlock) {
      const int loc = (int)(sk[q] & 0xFFFu);
      slot_row[slot0 + q] = loc < nrow ? row0 + (uint32_t)loc : kHitPad;
      if ((q & (kWave - 1)) == 0) {
        // a slice's width is its first (largest) count
        const uint32_t width =
            loc < nrow ? (uint32_t)(0xFFFFFFFFull - ((sk[q] >> 12) & 0xFFFFFFFFull)) : 0u;
        sbase[(slot0 + q) / kWave] = (uint64_t)width * kWave;
      }
    }
  } else {
    CT z;
    z.x = RT(0);
    z.y = RT(0);
#pragma unroll
    for (int r = 0; r < ROWS; ++r)
      if (valid[r]) rlen[row0 + wslot + r * kWave + lane] = cnt[r];
    __syncthreads();  // every row's count is written before the padding reads them
    // padding of every slot of the chunk: rows past their hits, empty slots whole
    for (int q = tid; q < P; q += kBlock) {
      const int64_t sid = (slot0 + q) / kWave;
      const uint32_t width = (uint32_t)((sbase[sid + 1] - sbase[sid]) / kWave);
      if (width == 0) continue;
      const uint32_t rr = slot_row[slot0 + q];
      const uint32_t k0 = rr == kHitPad ? 0u : rlen[rr];
      const uint64_t e0 = sbase[sid] + (uint64_t)(q % kWave);
      for (uint32_t k = k0; k < width; ++k) {
        col[e0 + (uint64_t)k * kWave] = kHitPad;
        hcoef[e0 + (uint64_t)k * kWave] = z;
      }
    }
  }
}

// In-place exclusive scan of a[0..m) (uint64) by one workgroup.
__global__ void __launch_bounds__(1024) scan_u64_kernel(uint64_t* __restrict__ a, int64_t m) {
  __shared__ uint64_t part[1024];
  const int tid = threadIdx.x;
  const int64_t per = (m + 1023) / 1024;
  const int64_t lo = min(m, (int64_t)tid * per);
  const int64_t hi = min(m, lo + per);
  uint64_t s = 0;
  for (int64_t i = lo; i < hi; ++i) s += a[i];
  part[tid] = s;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const uint64_t t = tid >= off ? part[tid - off] : 0ull;
    __syncthreads();
    part[tid] += t;
    __syncthreads();
  }
  uint64_t run = part[tid] - s;
  for (int64_t i = lo; i < hi; ++i) {
    const uint64_t t = a[i];
    a[i] = run;
    run += t;
  }
}

__device__ inline double2 to_d2(float2 a) { return make_double2(a.x, a.y); }
__device__ inline double2 to_d2(double2 a) { return a; }
template <typename VT> __device__ inline VT from_d2(double2 a);
template <> __device__ inline float2 from_d2<float2>(double2 a) { return make_float2((float)a.x, (float)a.y); }
template <> __device__ inline double2 from_d2<double2>(double2 a) { return a; }

// out = diag v + the hit table's product. A wave takes S slices and walks
// their rows' entries side by side, so that their chains of dependent loads
// (slice header, row, entries, gathers) overlap: S = 4 pays where rows have
// few hits (the product is then the per-row streams, latency-bound at one
// slice per wave), S = 1 where the gathers dominate (more waves in flight).
// Every load is issued unconditionally from a valid address (a lane past its
// row's hits re-reads its last entry, a row without hits entry 0) and its
// product is masked, which keeps the loads free of branches. FP64 sums of
// exact products (complex64 operands are widened before multiplying), each
// row's in group order.
// In a batch, the slices of an instance whose flag inst_skip is set are left
// alone (slices_per_item slices per chunk, a multiple of S).
// DIAG: the rows' diagonal products enter (diag and v's own entries read);
// without it the product off the diagonal.
template <typename VT, int S, bool DIAG>
__global__ void __launch_bounds__(kBlock)
hit_matvec_kernel(int64_t n_slices, const uint32_t* __restrict__ slot_row,
                  const uint64_t* __restrict__ sbase, const uint32_t* __restrict__ col,
                  const typename CoefT<VT>::type* __restrict__ hcoef, const VT* __restrict__ diag,
                  const uint32_t* __restrict__ rlen, const VT* __restrict__ v,
                  VT* __restrict__ out, const int32_t* __restrict__ skip,
                  int32_t* __restrict__ counter, const uint16_t* __restrict__ item_inst,
                  const int32_t* __restrict__ inst_skip, int slices_per_item) {
  if (skip && skip[0] != 0) return;
  if (counter && blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(counter, 1);
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t sid0 = ((int64_t)blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave) * S;
  if (sid0 >= n_slices) return;
  if (inst_skip && inst_skip[item_inst[sid0 / slices_per_item]] != 0) return;
  if constexpr (S == 1) {
    // one slice: each lane walks its own row's hits, several entries per step
    const uint32_t slot = slot_row[sid0 * kWave + lane];
    if (slot == kHitPad) return;
    const int64_t row = slot;
    const uint64_t b0 = sbase[sid0] + lane;
    const int len = (int)rlen[row];
    double sr = 0.0, si = 0.0;
    if constexpr (DIAG) {
      const double2 dd = to_d2(diag[row]), xd = to_d2(v[row]);
      sr = dd.x * xd.x - dd.y * xd.y;
      si = dd.x * xd.y + dd.y * xd.x;
    }
    // hits per step: more gathers in flight for complex64 (1.75 vs 1.79 ms on the
    // 28-qubit chain); complex128 is faster at 4 (2.66 vs 2.74 ms)
    constexpr int U = sizeof(VT) == 8 ? 8 : 4;
    int k = 0;
    for (; k + U <= len; k += U) {
      uint32_t j[U];
      double2 c[U], x[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const uint64_t e = b0 + (uint64_t)(k + u) * kWave;
        j[u] = col[e];
        c[u] = to_d2(hcoef[e]);
      }
#pragma unroll
      for (int u = 0; u < U; ++u) x[u] = to_d2(v[j[u]]);
#pragma unroll
      for (int u = 0; u < U; ++u) {
        sr = fma(c[u].x, x[u].x, fma(-c[u].y, x[u].y, sr));
        si = fma(c[u].x, x[u].y, fma(c[u].y, x[u].x, si));
      }
    }
    for (; k < len; ++k) {
      const uint64_t e = b0 + (uint64_t)k * kWave;
      const double2 c = to_d2(hcoef[e]);
      const double2 x = to_d2(v[col[e]]);
      sr = fma(c.x, x.x, fma(-c.y, x.y, sr));
      si = fma(c.x, x.y, fma(c.y, x.x, si));
    }
    out[row] = from_d2<VT>(make_double2(sr, si));
    return;
  }
  bool valid[S];
  int64_t row[S];
  uint64_t b0[S];
  int len[S];
  double sr[S], si[S];
#pragma unroll
  for (int u = 0; u < S; ++u) {
    const int64_t sid = min(sid0 + u, n_slices - 1);
    const uint32_t slot = slot_row[sid * kWave + lane];
    valid[u] = sid0 + u < n_slices && slot != kHitPad;
    row[u] = valid[u] ? (int64_t)slot : 0;
    b0[u] = sbase[sid] + lane;
  }
  int kmax = 0;
#pragma unroll
  for (int u = 0; u < S; ++u) {
    len[u] = valid[u] ? (int)rlen[row[u]] : 0;
    kmax = max(kmax, len[u]);
    sr[u] = si[u] = 0.0;
    if constexpr (DIAG) {
      const double2 dd = to_d2(diag[row[u]]), xd = to_d2(v[row[u]]);
      sr[u] = valid[u] ? dd.x * xd.x - dd.y * xd.y : 0.0;
      si[u] = valid[u] ? dd.x * xd.y + dd.y * xd.x : 0.0;
    }
  }
  for (int k = 0; k < kmax; ++k) {
    uint32_t j[S];
    double2 c[S];
#pragma unroll
    for (int u = 0; u < S; ++u) {
      const uint64_t e = len[u] > 0 ? b0[u] + (uint64_t)min(k, len[u] - 1) * kWave : 0;
      j[u] = col[e];
      c[u] = to_d2(hcoef[e]);
    }
#pragma unroll
    for (int u = 0; u < S; ++u) {
      double2 x = to_d2(v[len[u] > 0 ? j[u] : (uint32_t)row[u]]);
      if (k >= len[u]) x = make_double2(0.0, 0.0);
      sr[u] = fma(c[u].x, x.x, fma(-c[u].y, x.y, sr[u]));
      si[u] = fma(c[u].x, x.y, fma(c[u].y, x.x, si[u]));
    }
  }
#pragma unroll
  for (int u = 0; u < S; ++u)
    if (valid[u]) out[row[u]] = from_d2<VT>(make_double2(sr[u], si[u]));
}

// 32-bit words of class mask per chunk: 4 bits (one per wave) per class.
static_assert(kBlock / kWave == 4, "the class mask holds four waves per class");
inline int class_mask_stride(int n_classes) { return (n_classes + 7) / 8; }


// Tables of one hit-table pass (a single instance or a batch).
struct HitTables {
  const uint64_t* keys;
  const uint32_t* bstart;
  const uint16_t* fdir;
  const uint32_t* fbits;
  const uint32_t* items;
  const uint64_t* xmask;
  const int32_t* grp_off;
  const uint64_t* zmask;
  const void* coef;
  const uint32_t* class_hi;
  const int32_t* class_off;
  const sqdk_inst* descs;      // NULL: single instance
  const uint16_t* item_inst;
  sqdk_inst single;
  int n_items, chunk_rows, max_bucket, max_fine_bits, narrow;
  int dense_fine_bits;         // dense-lookup chunks: their largest fine_bits
  bool has_dense, has_search;
  int sort_log2;               // rows of a sort window: 2^sort_log2 (6 .. log2 chunk rows)
  uint32_t* cmask;             // class mask (cstride + 1 words per chunk), or NULL
  int cstride;
  uint32_t* rowhits;           // kShortHits source rows per row (with cmask)
};

// Build workspace: the class masks and wave flags (n_items x (stride + 1)
// words), then, 256-byte aligned, kShortHits source rows per row.
struct BuildWork {
  size_t mask_bytes, bytes;
};

inline BuildWork build_work_layout(int64_t n_items, int n_classes, int64_t n) {
  BuildWork b;
  b.mask_bytes = (((size_t)n_items * (class_mask_stride(n_classes) + 1) * 4 + 255) / 256) * 256;
  b.bytes = b.mask_bytes + (size_t)n * kShortHits * 4;
  return b;
}

// Points the tables at a caller's build workspace (or at none).
inline int attach_build_work(HitTables& t, int n_classes, int64_t n, const void* work,
                             size_t work_bytes) {
  t.cstride = class_mask_stride(n_classes);
  t.cmask = nullptr;
  t.rowhits = nullptr;
  if (!work) return SQDK_OK;
  const BuildWork b = build_work_layout(t.n_items, n_classes, n);
  if (work_bytes < b.bytes)
    return fail(SQDK_ENOSPC, "build workspace of %zu bytes, %zu needed", work_bytes, b.bytes);
  t.cmask = reinterpret_cast<uint32_t*>(const_cast<void*>(work));
  t.rowhits = reinterpret_cast<uint32_t*>(static_cast<char*>(const_cast<void*>(work)) +
                                          b.mask_bytes);
  return SQDK_OK;
}

template <typename VT, typename KT, int ROWS, bool DENSE, int MODE>
int launch_hit_table(const HitTables& t, uint32_t* slot_row, uint64_t* sbase, uint32_t* col,
                     void* hcoef, void* diag, uint32_t* rlen, hipStream_t s) {
  auto kern = hit_table_kernel<VT, KT, ROWS, DENSE, MODE>;
  SQDK_TRY(check_lds_base(reinterpret_cast<const void*>(kern), "hit_table_kernel"));
  const int dmax = DENSE ? t.dense_fine_bits : t.max_fine_bits;
  // (the chunk's sort keys, or its slot map, reuse the LDS after the classes)
  const size_t lds = std::max<size_t>(
      DENSE ? ((size_t)8 << dmax) : lds_layout(dmax, t.max_bucket, sizeof(KT), 0).total,
      (size_t)8 * ROWS * kBlock);
  if (lds > 160 * 1024) return fail(SQDK_ERANGE, "LDS use of %zu bytes exceeds 160 KiB", lds);
  if (lds > 65536)
    SQDK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  kern<<<(unsigned)t.n_items, kBlock, lds, s>>>(
      t.keys, t.bstart, t.fdir, t.fbits, t.items, t.xmask, t.grp_off, t.zmask,
      reinterpret_cast<const typename CoefT<VT>::type*>(t.coef), t.class_hi, t.class_off,
      t.descs, t.item_inst, t.single, dmax, t.max_bucket, t.sort_log2, slot_row, sbase, col,
      reinterpret_cast<typename CoefT<VT>::type*>(hcoef), reinterpret_cast<VT*>(diag), rlen,
      t.cmask, t.cstride, t.rowhits);
  SQDK_LAUNCHED("hit_table_kernel");
  return SQDK_OK;
}

template <typename VT, int MODE, int R>
int hit_table_rows(const HitTables& t0, uint32_t* slot_row, uint64_t* sbase, uint32_t* col,
                   void* hcoef, void* diag, uint32_t* rlen, hipStream_t s) {
  HitTables t = t0;
  t.sort_log2 = kHitSortLog2;
  if (t.has_dense)
    SQDK_TRY((launch_hit_table<VT, uint32_t, R, true, MODE>(t, slot_row, sbase, col, hcoef, diag,
                                                            rlen, s)));
  if (t.has_search) {
    if (t.narrow)
      return launch_hit_table<VT, uint32_t, R, false, MODE>(t, slot_row, sbase, col, hcoef, diag,
                                                            rlen, s);
    return launch_hit_table<VT, uint64_t, R, false, MODE>(t, slot_row, sbase, col, hcoef, diag,
                                                          rlen, s);
  }
  return SQDK_OK;
}

template <typename VT, int MODE>
int hit_table_dispatch(const HitTables& t, uint32_t* slot_row, uint64_t* sbase, uint32_t* col,
                       void* hcoef, void* diag, uint32_t* rlen, hipStream_t s) {
  switch (t.chunk_rows / kBlock) {
    case 4: return hit_table_rows<VT, MODE, 4>(t, slot_row, sbase, col, hcoef, diag, rlen, s);
    case 8: return hit_table_rows<VT, MODE, 8>(t, slot_row, sbase, col, hcoef, diag, rlen, s);
    case 16: return hit_table_rows<VT, MODE, 16>(t, slot_row, sbase, col, hcoef, diag, rlen, s);
    default:
      return fail(SQDK_EINVAL, "chunk of %d rows: the hit table needs 1024, 2048 or 4096",
                  t.chunk_rows);
  }
}

HitTables hit_tables_single(const sqdk_ham* h, const sqdk_space* sp, const void* coef) {
  HitTables t{};
  t.keys = sp->keys;
  t.bstart = sp->bstart;
  t.fdir = sp->fdir;
  t.fbits = sp->fbits;
  t.items = sp->items;
  t.xmask = h->xmask;
  t.grp_off = h->grp_off;
  t.zmask = h->zmask;
  t.coef = coef;
  t.class_hi = h->class_hi;
  t.class_off = h->class_off;
  t.descs = nullptr;
  t.item_inst = nullptr;
  t.single.bstart_base = 0;
  t.single.fdir_base = 0;
  t.single.class_base = 0;
  t.single.n_classes = h->n_classes;
  t.single.diag_t0 = h->diag_t0;
  t.single.diag_t1 = h->diag_t1;
  t.single.bucket_bits = sp->bucket_bits;
  t.single.fine_bits = sp->fine_bits;
  t.single.fbits_base = 0;
  t.n_items = sp->n_items;
  t.chunk_rows = sp->chunk_rows;
  t.max_bucket = sp->max_bucket;
  t.max_fine_bits = sp->fine_bits;
  t.dense_fine_bits = sp->fine_bits;
  t.narrow = sp->bucket_bits < 32;
  t.has_dense = sp->fbits != nullptr;
  t.has_search = sp->fbits == nullptr;
  return t;
}

HitTables hit_tables_batch(const sqdk_batch* b, const void* coef) {
  HitTables t{};
  t.keys = b->keys;
  t.bstart = b->bstart;
  t.fdir = b->fdir;
  t.fbits = b->fbits;
  t.items = b->items;
  t.xmask = b->xmask;
  t.grp_off = b->grp_off;
  t.zmask = b->zmask;
  t.coef = coef;
  t.class_hi = b->class_hi;
  t.class_off = b->class_off;
  t.descs = b->inst;
  t.item_inst = b->item_inst;
  memset(&t.single, 0, sizeof(t.single));
  t.n_items = b->n_items;
  t.chunk_rows = b->chunk_rows;
  t.max_bucket = b->max_bucket;
  t.max_fine_bits = b->max_fine_bits;
  t.dense_fine_bits = b->dense_fine_bits;
  t.narrow = b->max_bucket_bits < 32;
  t.has_dense = b->fbits != nullptr;
  t.has_search = true;   // a launch with no searched chunk returns at once
  return t;
}

int check_hit_space(const sqdk_ham* ham, const sqdk_space* space) {
  SQDK_TRY(check_space_public(space));
  SQDK_REQUIRE(ham != nullptr, "ham is NULL");
  SQDK_REQUIRE(space->fdir && space->items, "space lacks fine directory or items");
  SQDK_REQUIRE(ham->bucket_bits == space->bucket_bits && ham->num_qubits == space->num_qubits,
               "ham and space do not match");
  SQDK_REQUIRE(ham->n_classes == 0 || (ham->xmask && ham->grp_off && ham->class_hi && ham->class_off),
               "ham has NULL group tables");
  SQDK_REQUIRE(space->max_bucket >= 1 && space->max_bucket < 65536, "max_bucket %d out of range",
               space->max_bucket);
  SQDK_REQUIRE(space->fine_bits >= 0 && space->fine_bits <= 15 &&
                   space->fine_bits <= space->bucket_bits,
               "fine_bits %d invalid", space->fine_bits);
  SQDK_REQUIRE(!space->fbits || (space->fine_bits == space->bucket_bits - 5 && space->bucket_bits <= 20),
               "dense lookup needs fine_bits = bucket_bits - 5 <= 15");
  SQDK_REQUIRE(space->chunk_rows == 1024 || space->chunk_rows == 2048 || space->chunk_rows == 4096,
               "chunk of %d rows: the hit table needs 1024, 2048 or 4096", space->chunk_rows);
  return SQDK_OK;
}

}  // namespace

size_t matvec_work_bytes(const sqdk_space* space, const sqdk_tune* tune, int32_t precision) {
  const int split = (tune && tune->group_split > 1) ? tune->group_split : 1;
  if (split <= 1 || !space) return 0;
  return (size_t)(split - 1) * (size_t)space->n * (precision == SQDK_C128 ? 16 : 8);
}

int matvec_impl(const sqdk_ham* ham, const sqdk_space* space, int32_t precision,
                const void* v, void* out, const sqdk_tune* tune, void* work, size_t work_bytes,
                hipStream_t s, const int32_t* skip, int32_t* counter) {
  SQDK_TRY(check_space_public(space));
  SQDK_REQUIRE(ham != nullptr, "ham is NULL");
  SQDK_REQUIRE(v && out, "NULL vector");
  SQDK_REQUIRE(space->fdir && space->items, "space lacks fine directory or items");
  SQDK_REQUIRE(ham->bucket_bits == space->bucket_bits,
               "ham classes were built for bucket_bits %d, space has %d", ham->bucket_bits,
               space->bucket_bits);
  SQDK_REQUIRE(ham->num_qubits == space->num_qubits, "qubit counts differ (%d vs %d)",
               ham->num_qubits, space->num_qubits);
  SQDK_REQUIRE(ham->n_classes == 0 || (ham->xmask && ham->grp_off && ham->class_hi && ham->class_off),
               "ham has NULL group tables");
  SQDK_REQUIRE(ham->n_terms == 0 || ham->zmask, "ham has NULL zmask");
  SQDK_REQUIRE(space->max_bucket >= 1 && space->max_bucket < 65536, "max_bucket %d out of range",
               space->max_bucket);
  SQDK_REQUIRE(space->bucket_bits <= 63, "bucket_bits %d too large", space->bucket_bits);
  SQDK_REQUIRE(space->fine_bits >= 0 && space->fine_bits <= 15 &&
                   space->fine_bits <= space->bucket_bits,
               "fine_bits %d invalid", space->fine_bits);
  if (space->n_items == 0) return SQDK_OK;
  int rows = 4, xcd = 1, stage_v = 0, split = 1, compact = 0;
  if (tune) {
    if (tune->rows_per_thread) rows = tune->rows_per_thread;
    if (tune->group_split > 1) split = tune->group_split;
    xcd = tune->xcd_remap;
    stage_v = tune->stage_v;
    compact = tune->compact != 0;
  }
  SQDK_REQUIRE(split <= 1024, "group_split %d too large", split);
  SQDK_REQUIRE(rows * kBlock == space->chunk_rows || (rows == 4 && rows * 512 == space->chunk_rows),
               "rows_per_thread %d does not match the space's chunk of %d rows", rows,
               space->chunk_rows);
  const size_t need = matvec_work_bytes(space, tune, precision);
  SQDK_REQUIRE(need == 0 || (work && work_bytes >= need),
               "matvec workspace of %zu bytes, %zu needed for group_split %d", work_bytes, need,
               split);
  if (precision == SQDK_C128) {
    SQDK_REQUIRE(ham->n_terms == 0 || ham->coef, "ham has no complex128 coefficients");
    return launch_vt<double2>(ham, space, ham->coef, v, out, rows, xcd, stage_v, split, compact,
                              work, s, skip, counter);
  }
  if (precision == SQDK_C64) {
    SQDK_REQUIRE(ham->n_terms == 0 || ham->coef32, "ham has no complex64 coefficients");
    return launch_vt<float2>(ham, space, ham->coef32, v, out, rows, xcd, stage_v, split, compact,
                             work, s, skip, counter);
  }
  return fail(SQDK_EINVAL, "unknown precision %d", precision);
}

int hit_matvec_impl(const sqdk_cache* c, int32_t precision, const void* v, void* out,
                    hipStream_t s, const int32_t* skip, int32_t* counter,
                    const uint16_t* item_inst, const int32_t* inst_skip, int slices_per_item,
                    bool offdiag) {
  SQDK_REQUIRE(c != nullptr, "cache is NULL");
  SQDK_REQUIRE(v && out, "NULL vector");
  SQDK_REQUIRE(c->precision == precision, "cache precision %d, vectors %d", c->precision,
               precision);
  SQDK_REQUIRE(c->n >= 1 && c->n < (1ll << 31) && c->n_slices >= 1, "cache sizes out of range");
  SQDK_REQUIRE(c->slot_row && c->slice_base && c->diag && c->row_len,
               "cache has NULL tables");
  SQDK_REQUIRE(c->slices_per_wave == 0 || c->slices_per_wave == 1 || c->slices_per_wave == 4,
               "slices_per_wave %d not 1 or 4", c->slices_per_wave);
  SQDK_REQUIRE(!inst_skip || (item_inst && slices_per_item % 4 == 0),
               "instance skipping needs the chunk of each slice");
  const int spw = c->slices_per_wave == 4 ? 4 : 1;
  const int64_t grid = ceil_div(c->n_slices, (kBlock / kWave) * spw);
  SQDK_REQUIRE(grid < (1ll << 31), "too many slices");
  auto launch = [&](auto kern, auto vt) {
    using VT = decltype(vt);
    kern<<<(unsigned)grid, kBlock, 0, s>>>(
        c->n_slices, c->slot_row, c->slice_base, c->col,
        reinterpret_cast<const typename CoefT<VT>::type*>(c->coef),
        offdiag ? nullptr : reinterpret_cast<const VT*>(c->diag), c->row_len,
        reinterpret_cast<const VT*>(v),
        reinterpret_cast<VT*>(out), skip, counter, item_inst, inst_skip,
        slices_per_item > 0 ? slices_per_item : 1);
  };
  if (precision == SQDK_C128) {
    if (spw == 4)
      offdiag ? launch(hit_matvec_kernel<double2, 4, false>, double2{})
              : launch(hit_matvec_kernel<double2, 4, true>, double2{});
    else
      offdiag ? launch(hit_matvec_kernel<double2, 1, false>, double2{})
              : launch(hit_matvec_kernel<double2, 1, true>, double2{});
  } else if (precision == SQDK_C64) {
    if (spw == 4)
      offdiag ? launch(hit_matvec_kernel<float2, 4, false>, float2{})
              : launch(hit_matvec_kernel<float2, 4, true>, float2{});
    else
      offdiag ? launch(hit_matvec_kernel<float2, 1, false>, float2{})
              : launch(hit_matvec_kernel<float2, 1, true>, float2{});
  } else {
    return fail(SQDK_EINVAL, "unknown precision %d", precision);
  }
  SQDK_LAUNCHED("hit_matvec_kernel");
  return SQDK_OK;
}

int batch_matvec_impl(const sqdk_batch* b, int32_t precision, const void* v, void* out,
                      const sqdk_tune* tune, void* work, size_t work_bytes, hipStream_t s,
                      const int32_t* skip, const int32_t* inst_skip, int32_t* counter) {
  SQDK_REQUIRE(b != nullptr, "batch is NULL");
  SQDK_REQUIRE(v && out, "NULL vector");
  SQDK_REQUIRE(b->keys && b->bstart && b->fdir && b->items && b->item_inst && b->inst,
               "batch has NULL subspace tables");
  SQDK_REQUIRE(b->n >= 1 && b->n < (1ll << 31), "batch size %lld out of range", (long long)b->n);
  SQDK_REQUIRE(b->n_inst >= 1 && b->n_inst <= 65535, "n_inst %d out of range", b->n_inst);
  SQDK_REQUIRE(b->max_bucket >= 1 && b->max_bucket < 65536, "max_bucket %d out of range",
               b->max_bucket);
  SQDK_REQUIRE(b->max_fine_bits >= 0 && b->max_fine_bits <= 15, "max_fine_bits %d invalid",
               b->max_fine_bits);
  SQDK_REQUIRE(b->max_bucket_bits >= 0 && b->max_bucket_bits <= 63, "max_bucket_bits %d invalid",
               b->max_bucket_bits);
  SQDK_REQUIRE(!b->fbits || (b->dense_fine_bits >= 0 && b->dense_fine_bits <= 15),
               "dense_fine_bits %d invalid", b->dense_fine_bits);
  if (b->n_items == 0) return SQDK_OK;
  int rows = 4, xcd = 1, stage_v = 0, split = 1, compact = 0;
  if (tune) {
    if (tune->rows_per_thread) rows = tune->rows_per_thread;
    if (tune->group_split > 1) split = tune->group_split;
    xcd = tune->xcd_remap;
    stage_v = tune->stage_v;
    compact = tune->compact != 0;
  }
  SQDK_REQUIRE(split <= 1024, "group_split %d too large", split);
  SQDK_REQUIRE(rows * kBlock == b->chunk_rows || (rows == 4 && rows * 512 == b->chunk_rows),
               "rows_per_thread %d does not match the batch's chunk of %d rows", rows,
               b->chunk_rows);
  const size_t need = split > 1 ? (size_t)(split - 1) * (size_t)b->n * (precision == SQDK_C128 ? 16 : 8) : 0;
  SQDK_REQUIRE(need == 0 || (work && work_bytes >= need),
               "matvec workspace of %zu bytes, %zu needed for group_split %d", work_bytes, need,
               split);
  if (precision == SQDK_C128)
    return launch_batch<double2>(b, b->coef, v, out, rows, xcd, stage_v, split, compact, work, s,
                                 skip, inst_skip, counter);
  if (precision == SQDK_C64)
    return launch_batch<float2>(b, b->coef32, v, out, rows, xcd, stage_v, split, compact, work,
                                s, skip, inst_skip, counter);
  return fail(SQDK_EINVAL, "unknown precision %d", precision);
}

}  // namespace sqdk

extern "C" int sqdk_batch_matvec_workspace(const sqdk_batch* batch, const sqdk_tune* tune,
                                           int32_t precision, size_t* bytes) {
  using namespace sqdk;
  SQDK_REQUIRE(batch && bytes, "NULL pointer argument");
  const int split = (tune && tune->group_split > 1) ? tune->group_split : 1;
  *bytes = split > 1 ? (size_t)(split - 1) * (size_t)batch->n * (precision == SQDK_C128 ? 16 : 8) : 0;
  return SQDK_OK;
}

extern "C" int sqdk_batch_matvec(const sqdk_batch* batch, int32_t precision, const void* v,
                                 void* out, const sqdk_tune* tune, void* work, size_t work_bytes,
                                 void* stream) {
  return sqdk::batch_matvec_impl(batch, precision, v, out, tune, work, work_bytes,
                                 sqdk::as_stream(stream), nullptr, nullptr, nullptr);
}

extern "C" int sqdk_matvec_workspace(const sqdk_space* space, const sqdk_tune* tune,
                                     int32_t precision, size_t* bytes) {
  using namespace sqdk;
  SQDK_REQUIRE(space && bytes, "NULL pointer argument");
  SQDK_REQUIRE(precision == SQDK_C128 || precision == SQDK_C64, "unknown precision %d", precision);
  *bytes = sqdk::matvec_work_bytes(space, tune, precision);
  return SQDK_OK;
}

extern "C" int sqdk_matvec(const sqdk_ham* ham, const sqdk_space* space, int32_t precision,
                           const void* v, void* out, const sqdk_tune* tune, void* work,
                           size_t work_bytes, void* stream) {
  return sqdk::matvec_impl(ham, space, precision, v, out, tune, work, work_bytes,
                           sqdk::as_stream(stream));
}

extern "C" int sqdk_cache_slices(const sqdk_space* space, int64_t* n_slices) {
  using namespace sqdk;
  SQDK_REQUIRE(space && n_slices, "NULL pointer argument");
  SQDK_REQUIRE(space->chunk_rows == 1024 || space->chunk_rows == 2048 || space->chunk_rows == 4096,
               "chunk of %d rows: the hit table needs 1024, 2048 or 4096", space->chunk_rows);
  *n_slices = (int64_t)space->n_items * (space->chunk_rows / kWave);
  return SQDK_OK;
}

extern "C" int sqdk_cache_workspace(const sqdk_ham* ham, const sqdk_space* space,
                                    size_t* bytes) {
  using namespace sqdk;
  SQDK_REQUIRE(ham && space && bytes, "NULL pointer argument");
  SQDK_REQUIRE(space->n_items >= 0 && space->n >= 0 && ham->n_classes >= 0,
               "negative size in ham or space");
  *bytes = build_work_layout(space->n_items, ham->n_classes, space->n).bytes;
  return SQDK_OK;
}

extern "C" int sqdk_cache_count(const sqdk_ham* ham, const sqdk_space* space,
                                uint32_t* slot_row, uint64_t* slice_base, uint32_t* row_len,
                                void* work, size_t work_bytes, void* stream) {
  using namespace sqdk;
  SQDK_TRY(check_hit_space(ham, space));
  SQDK_REQUIRE(slot_row && slice_base && row_len, "NULL pointer argument");
  SQDK_REQUIRE(space->n_items >= 1, "empty subspace");
  hipStream_t s = as_stream(stream);
  const int64_t ns = (int64_t)space->n_items * (space->chunk_rows / kWave);
  SQDK_HIP(hipMemsetAsync(slice_base + ns, 0, sizeof(uint64_t), s));
  HitTables t = hit_tables_single(ham, space, ham->coef32);
  SQDK_TRY(attach_build_work(t, ham->n_classes, space->n, work, work_bytes));
  SQDK_TRY((hit_table_dispatch<float2, 0>(t, slot_row, slice_base, nullptr, nullptr, nullptr,
                                          row_len, s)));
  scan_u64_kernel<<<1, 1024, 0, s>>>(slice_base, ns + 1);
  SQDK_LAUNCHED("scan_u64_kernel");
  return SQDK_OK;
}

extern "C" int sqdk_cache_fill(const sqdk_ham* ham, const sqdk_space* space, int32_t precision,
                               const uint64_t* slice_base, const uint32_t* slot_row, uint32_t* col,
                               void* coef, void* diag, uint32_t* row_len, const void* work,
                               size_t work_bytes, void* stream) {
  using namespace sqdk;
  SQDK_TRY(check_hit_space(ham, space));
  SQDK_REQUIRE(slice_base && slot_row && col && coef && diag && row_len,
               "NULL pointer argument");
  hipStream_t s = as_stream(stream);
  uint64_t* sb = const_cast<uint64_t*>(slice_base);
  uint32_t* sr = const_cast<uint32_t*>(slot_row);
  if (precision != SQDK_C128 && precision != SQDK_C64)
    return fail(SQDK_EINVAL, "unknown precision %d", precision);
  const bool wide = precision == SQDK_C128;
  SQDK_REQUIRE(ham->n_terms == 0 || (wide ? (const void*)ham->coef : (const void*)ham->coef32),
               "ham has no %s coefficients", wide ? "complex128" : "complex64");
  HitTables t = hit_tables_single(ham, space, wide ? (const void*)ham->coef
                                                   : (const void*)ham->coef32);
  SQDK_TRY(attach_build_work(t, ham->n_classes, space->n, work, work_bytes));
  return wide ? hit_table_dispatch<double2, 1>(t, sr, sb, col, coef, diag, row_len, s)
              : hit_table_dispatch<float2, 1>(t, sr, sb, col, coef, diag, row_len, s);
}

extern "C" int sqdk_cache_matvec(const sqdk_cache* cache, const void* v, void* out,
                                 void* stream) {
  using namespace sqdk;
  SQDK_REQUIRE(cache != nullptr, "cache is NULL");
  return hit_matvec_impl(cache, cache->precision, v, out, as_stream(stream), nullptr, nullptr);
}

extern "C" int sqdk_batch_cache_slices(const sqdk_batch* batch, int64_t* n_slices) {
  using namespace sqdk;
  SQDK_REQUIRE(batch && n_slices, "NULL pointer argument");
  SQDK_REQUIRE(batch->chunk_rows == 1024 || batch->chunk_rows == 2048 || batch->chunk_rows == 4096,
               "chunk of %d rows: the hit table needs 1024, 2048 or 4096", batch->chunk_rows);
  *n_slices = (int64_t)batch->n_items * (batch->chunk_rows / kWave);
  return SQDK_OK;
}

extern "C" int sqdk_batch_cache_workspace(const sqdk_batch* batch, int32_t max_classes,
                                          size_t* bytes) {
  using namespace sqdk;
  SQDK_REQUIRE(batch && bytes, "NULL pointer argument");
  SQDK_REQUIRE(batch->n_items >= 0 && batch->n >= 0 && max_classes >= 0,
               "negative size in batch or max_classes");
  *bytes = build_work_layout(batch->n_items, max_classes, batch->n).bytes;
  return SQDK_OK;
}

extern "C" int sqdk_batch_cache_count(const sqdk_batch* batch, uint32_t* slot_row,
                                      uint64_t* slice_base, uint32_t* row_len,
                                      int32_t max_classes, void* work, size_t work_bytes,
                                      void* stream) {
  using namespace sqdk;
  SQDK_REQUIRE(batch && slot_row && slice_base && row_len, "NULL pointer argument");
  SQDK_REQUIRE(batch->keys && batch->bstart && batch->fdir && batch->items && batch->item_inst &&
                   batch->inst, "batch has NULL subspace tables");
  SQDK_REQUIRE(batch->n_items >= 1, "empty batch");
  SQDK_REQUIRE(max_classes >= 0, "negative max_classes %d", max_classes);
  int64_t ns = 0;
  SQDK_TRY(sqdk_batch_cache_slices(batch, &ns));
  hipStream_t s = as_stream(stream);
  SQDK_HIP(hipMemsetAsync(slice_base + ns, 0, sizeof(uint64_t), s));
  HitTables t = hit_tables_batch(batch, batch->coef32);
  SQDK_TRY(attach_build_work(t, max_classes, batch->n, work, work_bytes));
  SQDK_TRY((hit_table_dispatch<float2, 0>(t, slot_row, slice_base, nullptr, nullptr, nullptr,
                                          row_len, s)));
  scan_u64_kernel<<<1, 1024, 0, s>>>(slice_base, ns + 1);
  SQDK_LAUNCHED("scan_u64_kernel");
  return SQDK_OK;
}

extern "C" int sqdk_batch_cache_fill(const sqdk_batch* batch, int32_t precision,
                                     const uint64_t* slice_base, const uint32_t* slot_row,
                                     uint32_t* col, void* coef, void* diag, uint32_t* row_len,
                                     int32_t max_classes, const void* work, size_t work_bytes,
                                     void* stream) {
  using namespace sqdk;
  SQDK_REQUIRE(batch && slice_base && slot_row && col && coef && diag && row_len,
               "NULL pointer argument");
  SQDK_REQUIRE(max_classes >= 0, "negative max_classes %d", max_classes);
  hipStream_t s = as_stream(stream);
  uint64_t* sb = const_cast<uint64_t*>(slice_base);
  uint32_t* sr = const_cast<uint32_t*>(slot_row);
  if (precision != SQDK_C128 && precision != SQDK_C64)
    return fail(SQDK_EINVAL, "unknown precision %d", precision);
  const bool wide = precision == SQDK_C128;
  HitTables t = hit_tables_batch(batch, wide ? (const void*)batch->coef
                                             : (const void*)batch->coef32);
  SQDK_TRY(attach_build_work(t, max_classes, batch->n, work, work_bytes));
  return wide ? hit_table_dispatch<double2, 1>(t, sr, sb, col, coef, diag, row_len, s)
              : hit_table_dispatch<float2, 1>(t, sr, sb, col, coef, diag, row_len, s);
}
