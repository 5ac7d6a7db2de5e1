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
#include "internal.hpp"

#include <cstring>

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
struct LdsLayout {
  size_t pairs, occ, keys, vals, total;
};

__host__ __device__ inline LdsLayout lds_layout(int fine_bits, int max_bucket, size_t key_bytes, size_t val_bytes) {
  LdsLayout l;
  l.pairs = 0;
  l.occ = ((size_t)4 << fine_bits);
  l.keys = l.occ + 16 * sizeof(int);
  l.vals = (l.keys + key_bytes * ((size_t)max_bucket + 1) + 15) & ~(size_t)15;
  l.total = l.vals + val_bytes * (size_t)max_bucket;
  return l;
}

template <typename VT> constexpr bool kGatherFirst = sizeof(VT) == 8;

// sum_k c_k (-1)^{popcount(z_k & s_r)} for every row r of a thread, terms in
// index order (the order of group_coef, so results are identical). The term
// tables are read once per term for all rows.
// NARROW: every row key of the chunk is below 2^32, so the parity needs only
// the low words.
template <typename CT, typename KT, int ROWS, bool NARROW>
__device__ inline void rows_coef_n(const uint64_t* __restrict__ zmask, const CT* __restrict__ coef,
                                   int t0, int t1, uint64_t bucket_base, const KT (&sl)[ROWS],
                                   decltype(CT::x) (&cr)[ROWS], decltype(CT::x) (&ci)[ROWS]) {
#pragma unroll
  for (int r = 0; r < ROWS; ++r) {
    cr[r] = 0;
    ci[r] = 0;
  }
  for (int k = t0; k < t1; ++k) {
    const uint64_t z = zmask[k];
    const CT c = coef[k];
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
      const int pc = NARROW ? __popc((uint32_t)z & ((uint32_t)bucket_base | (uint32_t)sl[r]))
                            : __popcll(z & (bucket_base | (uint64_t)sl[r]));
      const bool odd = (pc & 1) != 0;
      cr[r] += odd ? -c.x : c.x;
      ci[r] += odd ? -c.y : c.y;
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

// Occupancy: at 4 rows per thread the kernel must keep 5 waves per SIMD
// (at most 96 VGPRs); the complex128 variant needs the hint to get there.
template <typename VT, typename KT, int ROWS, bool STAGE_V, bool DENSE, int NT>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(ROWS == 4 ? 5 : 1)))
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
  VT* lv = reinterpret_cast<VT*>(smem + lay.vals);
  // Dense lookup: (occupancy bitmap, first-key offset) of every 32-value
  // sub-bucket of the staged target bucket.
  uint2* ldir = reinterpret_cast<uint2*>(smem);

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
#pragma unroll
  for (int r = 0; r < ROWS; ++r) {
    const uint32_t i = row0 + tid + r * NT;
    valid[r] = i < row1;
    sl[r] = valid[r] ? (KT)(keys[i] & lowmask) : (KT)0;
  }

  // Waves of a short chunk that hold no row skip the per-group work.
  const bool wave_active = __any(valid[0]);

  // X = 0 group: the diagonal (first share only), one pass over its terms
  // for all rows of the thread.
  if (gs == 0 && D.diag_t1 > D.diag_t0 && wave_active) {
    RT cr[ROWS], ci[ROWS];
    rows_coef<CT, KT, ROWS>(zmask, coef, D.diag_t0, D.diag_t1, bucket_base, narrow, sl, cr, ci);
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
      if (valid[r]) acc[r].fma(cr[r], ci[r], v[row0 + tid + r * NT]);
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
      const uint64_t fbase = q << d;
      for (int e = tid; e < ndir; e += NT) ldir[e] = make_uint2(fbits[fbase + e], fdr[fbase + e]);
      __syncthreads();
      if (!wave_active) continue;  // the barriers above are still reached by every wave
      const VT* __restrict__ vq = v + qs;
      for (int g = gfirst; g < g1; g += a.split) {
        const uint32_t xl = (uint32_t)(xmask[g] & lowmask);
        int idx[ROWS];
        bool any_hit = false;
#pragma unroll
        for (int r = 0; r < ROWS; ++r) {
          // sub-bucket of 32 values: one LDS read, a bit test and a rank
          const uint32_t t = (uint32_t)sl[r] ^ xl;
          const uint2 pr = ldir[t >> 5];
          const uint32_t b = t & 31u;
          const bool hit = valid[r] && ((pr.x >> b) & 1u);
          idx[r] = hit ? (int)(pr.y + __popc(pr.x & ((1u << b) - 1u))) : -1;
          any_hit |= hit;
        }
        if (!__any(any_hit)) continue;
        // (the dense path has registers to spare: gather first in both precisions)
        VT vv[ROWS];
#pragma unroll
        for (int r = 0; r < ROWS; ++r) vv[r] = vq[max(idx[r], 0)];
        RT cr[ROWS], ci[ROWS];
        rows_coef<CT, KT, ROWS>(zmask, coef, grp_off[g], grp_off[g + 1], bucket_base, narrow, sl, cr, ci);
#pragma unroll
        for (int r = 0; r < ROWS; ++r) {
          if (idx[r] >= 0) acc[r].fma(cr[r], ci[r], vv[r]);
        }
      }
      (void)qn;
    } else {
      __syncthreads();  // the previous class is done with the LDS tables
      for (int i = tid; i < qn; i += NT) lkeys[i] = (KT)(keys[qs + i] & lowmask);
      if (tid == 0) lkeys[qn] = (KT)0;
      if (STAGE_V) {
        for (int i = tid; i < qn; i += NT) lv[i] = v[qs + i];
      }
      // Sub-bucket ranges as (lo | hi << 16) and the largest occupancy.
      const uint64_t fbase = q << d;
      int occ = 0;
      for (int e = tid; e < ndir; e += NT) {
        const uint32_t lo = fdr[fbase + e];
        const uint32_t hi = (e + 1 < ndir) ? (uint32_t)fdr[fbase + e + 1] : (uint32_t)qn;
        lpair[e] = lo | (hi << 16);
        occ = max(occ, (int)(hi - lo));
      }
#pragma unroll
      for (int sh = kWave / 2; sh > 0; sh >>= 1) occ = max(occ, __shfl_xor(occ, sh));
      if ((tid & (kWave - 1)) == 0) s_occ[tid / kWave] = occ;
      __syncthreads();
      int maxocc = 0;
#pragma unroll
      for (int i = 0; i < NT / kWave; ++i) maxocc = max(maxocc, s_occ[i]);
      // Steps of the branch-free lower bound: ceil(log2(maxocc)).
      const int steps = maxocc <= 1 ? 0 : 32 - __clz(maxocc - 1);
      if (!wave_active) continue;  // the barriers above are still reached by every wave
      const VT* __restrict__ vq = v + qs;
      for (int g = gfirst; g < g1; g += a.split) {
        const KT xl = (KT)(xmask[g] & lowmask);
        KT tl[ROWS];
        int base[ROWS], n[ROWS];
        bool live[ROWS];
#pragma unroll
        for (int r = 0; r < ROWS; ++r) {
          tl[r] = sl[r] ^ xl;
          const uint32_t pr = lpair[(uint32_t)(tl[r] >> f)];
          base[r] = (int)(pr & 0xFFFFu);
          n[r] = (int)(pr >> 16) - base[r];
          live[r] = valid[r] && n[r] > 0;
        }
        // Branch-free search for the last key <= target: after ceil(log2(n))
        // halvings base is that key, so one equality test decides membership.
        // base stays in [lo, qn], and lkeys[qn] exists, so every read is in range.
        for (int st = 0; st < steps; ++st) {
#pragma unroll
          for (int r = 0; r < ROWS; ++r) {
            const int half = n[r] >> 1;
            base[r] = lkeys[base[r] + half] <= tl[r] ? base[r] + half : base[r];
            n[r] -= half;
          }
        }
        int idx[ROWS];
        bool any_hit = false;
#pragma unroll
        for (int r = 0; r < ROWS; ++r) {
          const KT kb = lkeys[base[r]];
          idx[r] = (live[r] && kb == tl[r]) ? base[r] : -1;
          any_hit |= idx[r] >= 0;
        }
        if (!__any(any_hit)) continue;
        // Hits: gather (a miss reads the slice's first element, in range and
        // shared by the wave) and one pass over the group's terms for all rows.
        // complex64 gathers first to hide the load behind the terms; complex128
        // gathers after them, which keeps 16 fewer VGPRs live (occupancy).
        VT vv[ROWS];
        if constexpr (kGatherFirst<VT>) {
#pragma unroll
          for (int r = 0; r < ROWS; ++r) vv[r] = STAGE_V ? lv[max(idx[r], 0)] : vq[max(idx[r], 0)];
        }
        RT cr[ROWS], ci[ROWS];
        rows_coef<CT, KT, ROWS>(zmask, coef, grp_off[g], grp_off[g + 1], bucket_base, narrow, sl, cr, ci);
#pragma unroll
        for (int r = 0; r < ROWS; ++r) {
          if constexpr (!kGatherFirst<VT>) vv[r] = STAGE_V ? lv[max(idx[r], 0)] : vq[max(idx[r], 0)];
          if (idx[r] >= 0) acc[r].fma(cr[r], ci[r], vv[r]);
        }
      }
    }
  }

  // Shares other than the first write partial sums for the reduction pass.
  VT* __restrict__ dst = gs == 0 ? out : reinterpret_cast<VT*>(a.partial) + (int64_t)(gs - 1) * a.n;
#pragma unroll
  for (int r = 0; r < ROWS; ++r) {
    if (valid[r]) dst[row0 + tid + r * NT] = acc[r].result();
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

template <typename VT, typename KT, int ROWS, bool SV, bool DENSE, int NT = kBlock>
int launch_rows(const MatvecTables<VT>& t, const MatvecParams& args, size_t lds, hipStream_t s) {
  // (t.descs / t.item_inst are NULL for a single instance)
  auto kern = bucket_join_kernel<VT, KT, ROWS, SV, DENSE, NT>;
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
  if (args.split > 1) {
    const unsigned g = (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(args.n, kBlock), 4096));
    split_reduce_kernel<VT><<<g, kBlock, 0, s>>>(args.n, args.split - 1,
                                                 reinterpret_cast<const VT*>(args.partial), t.out,
                                                 args.skip);
    SQDK_LAUNCHED("split_reduce_kernel");
  }
  return SQDK_OK;
}

template <typename VT, typename KT, bool SV, bool DENSE = false>
int launch_kt(const MatvecTables<VT>& t, const MatvecParams& args, int rows, size_t lds,
              hipStream_t s) {
  // 512-thread workgroups (2,048-row chunks) exist for 4 rows per thread only.
  if (args.threads == 512) {
    if (rows != 4) return fail(SQDK_EINVAL, "512-thread chunks need 4 rows per thread");
    return launch_rows<VT, KT, 4, SV, DENSE, 512>(t, args, lds, s);
  }
  switch (rows) {
    case 4: return launch_rows<VT, KT, 4, SV, DENSE>(t, args, lds, s);
    case 8: return launch_rows<VT, KT, 8, SV, DENSE>(t, args, lds, s);
    case 16: return launch_rows<VT, KT, 16, SV, DENSE>(t, args, lds, s);
    default: return fail(SQDK_EINVAL, "rows_per_thread %d unsupported (4, 8 or 16)", rows);
  }
}

template <typename VT>
int launch_common(const MatvecTables<VT>& t, MatvecParams& args, int narrow_keys, int rows,
                  int stage_v, hipStream_t s) {
  if (t.fbits) {
    // Dense lookup: only the (bitmap, offset) table of the target bucket is staged.
    if (stage_v) return fail(SQDK_EINVAL, "v staging is not available with the dense lookup");
    if (((size_t)8 << args.dmax) > 160 * 1024)
      return fail(SQDK_ERANGE, "dense lookup table of 2^%d entries exceeds 160 KiB of LDS", args.dmax);
    return launch_kt<VT, uint32_t, false, true>(t, args, rows, (size_t)8 << args.dmax, s);
  }
  const LdsLayout lay = lds_layout(args.dmax, args.max_bucket, narrow_keys ? 4 : 8,
                                   stage_v ? sizeof(VT) : 0);
  if (lay.total > 160 * 1024)
    return fail(SQDK_ERANGE, "LDS tables of %zu bytes exceed 160 KiB (largest bucket %d keys)",
                lay.total, args.max_bucket);
  if (stage_v)
    return narrow_keys ? launch_kt<VT, uint32_t, true>(t, args, rows, lay.total, s)
                       : launch_kt<VT, uint64_t, true>(t, args, rows, lay.total, s);
  return narrow_keys ? launch_kt<VT, uint32_t, false>(t, args, rows, lay.total, s)
                     : launch_kt<VT, uint64_t, false>(t, args, rows, lay.total, s);
}

template <typename VT>
int launch_vt(const sqdk_ham* h, const sqdk_space* sp, const void* coef, const void* v,
              void* out, int rows, int xcd, int stage_v, int split, void* work,
              hipStream_t s, const int32_t* skip, int32_t* counter) {
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
  return launch_common<VT>(t, args, sp->bucket_bits <= 32, rows, stage_v, s);
}

template <typename VT>
int launch_batch(const sqdk_batch* b, const void* coef, const void* v, void* out, int rows,
                 int xcd, int stage_v, int split, void* work, hipStream_t s,
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
  args.n = b->n;
  args.partial = work;
  memset(&args.single, 0, sizeof(args.single));
  return launch_common<VT>(t, args, b->max_bucket_bits <= 32, rows, stage_v, s);
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
  int rows = 4, xcd = 1, stage_v = 0, split = 1;
  if (tune) {
    if (tune->rows_per_thread) rows = tune->rows_per_thread;
    if (tune->group_split > 1) split = tune->group_split;
    xcd = tune->xcd_remap;
    stage_v = tune->stage_v;
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
    return launch_vt<double2>(ham, space, ham->coef, v, out, rows, xcd, stage_v, split, work, s,
                              skip, counter);
  }
  if (precision == SQDK_C64) {
    SQDK_REQUIRE(ham->n_terms == 0 || ham->coef32, "ham has no complex64 coefficients");
    return launch_vt<float2>(ham, space, ham->coef32, v, out, rows, xcd, stage_v, split, work, s,
                             skip, counter);
  }
  return fail(SQDK_EINVAL, "unknown precision %d", precision);
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
  if (b->n_items == 0) return SQDK_OK;
  int rows = 4, xcd = 1, stage_v = 0, split = 1;
  if (tune) {
    if (tune->rows_per_thread) rows = tune->rows_per_thread;
    if (tune->group_split > 1) split = tune->group_split;
    xcd = tune->xcd_remap;
    stage_v = tune->stage_v;
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
    return launch_batch<double2>(b, b->coef, v, out, rows, xcd, stage_v, split, work, s, skip,
                                 inst_skip, counter);
  if (precision == SQDK_C64)
    return launch_batch<float2>(b, b->coef32, v, out, rows, xcd, stage_v, split, work, s, skip,
                                inst_skip, counter);
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
