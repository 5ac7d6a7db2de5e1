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

namespace sqdk {

namespace {

constexpr int kBlock = 256;

template <typename VT> struct CoefT;
template <> struct CoefT<double2> { using type = double2; using real = double; };
template <> struct CoefT<float2> { using type = float2; using real = float; };

template <typename VT>
struct MatvecArgs {
  const uint64_t* __restrict__ keys;
  const uint32_t* __restrict__ bstart;
  const uint16_t* __restrict__ fdir;
  const uint32_t* __restrict__ items;
  const uint64_t* __restrict__ pc_mask;
  const uint64_t* __restrict__ xmask;
  const int32_t* __restrict__ grp_off;
  const uint64_t* __restrict__ zmask;
  const typename CoefT<VT>::type* __restrict__ coef;
  const uint32_t* __restrict__ class_hi;
  const int32_t* __restrict__ class_off;
  const VT* __restrict__ v;
  VT* __restrict__ out;
  const int32_t* __restrict__ skip;
  int32_t* __restrict__ counter;
  int32_t n_items;
  int32_t m;          // bucket bits
  int32_t d;          // fine bits
  int32_t n_classes;
  int32_t diag_t0, diag_t1;
  int32_t use_pc;
  int32_t xcd_remap;
};

// Workgroup i runs on XCD i % 8; give each XCD a contiguous range of items so
// chunks of one bucket, which read the same target buckets, share an L2.
__device__ inline int remap_xcd(int i, int g) {
  constexpr int kXcd = 8;
  const int per = g / kXcd, rem = g % kXcd;
  const int xcd = i % kXcd, local = i / kXcd;
  return xcd < rem ? xcd * (per + 1) + local : rem * (per + 1) + (xcd - rem) * per + local;
}

template <typename VT, typename KT, int ROWS>
__global__ void __launch_bounds__(kBlock)
bucket_join_kernel(MatvecArgs<VT> a) {
  using CT = typename CoefT<VT>::type;
  using RT = typename CoefT<VT>::real;
  extern __shared__ __align__(16) unsigned char smem[];
  const int tid = threadIdx.x;
  const int w = a.xcd_remap ? remap_xcd(blockIdx.x, gridDim.x) : (int)blockIdx.x;
  if (w >= a.n_items) return;
  if (a.skip && a.skip[0] != 0) return;
  if (a.counter && blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(a.counter, 1);

  const int m = a.m;
  const int d = a.d;
  const int f = m - d;
  const uint64_t lowmask = (m >= 64) ? ~0ull : ((1ull << m) - 1ull);
  const uint32_t dmask = (1u << d) - 1u;
  const int ndir = 1 << d;
  uint16_t* ldir = reinterpret_cast<uint16_t*>(smem);
  KT* lkeys = reinterpret_cast<KT*>(smem + (((ndir + 1) * 2 + 15) & ~15));

  const uint32_t row0 = a.items[w];
  const uint64_t bucket = a.keys[row0] >> m;
  const uint32_t bend = a.bstart[bucket + 1];
  const uint32_t row1 = min(bend, row0 + (uint32_t)(ROWS * kBlock));
  const uint64_t pc[2] = {a.pc_mask[0], a.pc_mask[1]};

  uint64_t s[ROWS];
  bool valid[ROWS];
  Acc<VT> acc[ROWS];
#pragma unroll
  for (int r = 0; r < ROWS; ++r) {
    const uint32_t i = row0 + tid + r * kBlock;
    valid[r] = i < row1;
    s[r] = valid[r] ? a.keys[i] : 0ull;
  }

  // X = 0 group: the diagonal.
  if (a.diag_t1 > a.diag_t0) {
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
      if (!valid[r]) continue;
      RT cr, ci;
      group_coef<CT>(a.zmask, a.coef, a.diag_t0, a.diag_t1, s[r], cr, ci);
      acc[r].fma(cr, ci, a.v[row0 + tid + r * kBlock]);
    }
  }

  for (int c = 0; c < a.n_classes; ++c) {
    const uint64_t q = bucket ^ (uint64_t)a.class_hi[c];
    const uint32_t qs = a.bstart[q];
    const uint32_t qe = a.bstart[q + 1];
    if (qs == qe) continue;  // empty target bucket: no group of the class can hit
    const int qn = (int)(qe - qs);
    __syncthreads();  // previous class is done with the LDS tables
    for (int i = tid; i < qn; i += kBlock) lkeys[i] = (KT)(a.keys[qs + i] & lowmask);
    for (int e = tid; e < ndir; e += kBlock) ldir[e] = a.fdir[(q << d) + e];
    if (tid == 0) ldir[ndir] = (uint16_t)qn;
    __syncthreads();
    const VT* __restrict__ vq = a.v + qs;
    const int g0 = a.class_off[c], g1 = a.class_off[c + 1];
    for (int g = g0; g < g1; ++g) {
      const uint64_t x = a.xmask[g];
      const int t0 = a.grp_off[g], t1 = a.grp_off[g + 1];
#pragma unroll
      for (int r = 0; r < ROWS; ++r) {
        if (!valid[r]) continue;
        const uint64_t t = s[r] ^ x;
        if (a.use_pc && !popcount_present(t, pc)) continue;
        const KT tl = (KT)(t & lowmask);
        const uint32_t e = (uint32_t)(t >> f) & dmask;
        int lo = ldir[e];
        int hi = ldir[e + 1];
        const int end = hi;
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          if (lkeys[mid] < tl) lo = mid + 1; else hi = mid;
        }
        if (lo < end && lkeys[lo] == tl) {
          RT cr, ci;
          group_coef<CT>(a.zmask, a.coef, t0, t1, s[r], cr, ci);
          acc[r].fma(cr, ci, vq[lo]);
        }
      }
    }
  }

#pragma unroll
  for (int r = 0; r < ROWS; ++r) {
    if (valid[r]) a.out[row0 + tid + r * kBlock] = acc[r].result();
  }
}

template <typename VT, typename KT, int ROWS>
int launch_rows(const MatvecArgs<VT>& args, size_t lds, hipStream_t s) {
  auto kern = bucket_join_kernel<VT, KT, ROWS>;
  if (lds > 65536) {
    SQDK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  }
  kern<<<args.n_items, kBlock, lds, s>>>(args);
  SQDK_LAUNCHED("bucket_join_kernel");
  return SQDK_OK;
}

template <typename VT, typename KT>
int launch_kt(const MatvecArgs<VT>& args, int rows, size_t lds, hipStream_t s) {
  switch (rows) {
    case 4: return launch_rows<VT, KT, 4>(args, lds, s);
    case 8: return launch_rows<VT, KT, 8>(args, lds, s);
    case 16: return launch_rows<VT, KT, 16>(args, lds, s);
    default: return fail(SQDK_EINVAL, "rows_per_thread %d unsupported (4, 8 or 16)", rows);
  }
}

template <typename VT>
int launch_vt(const sqdk_ham* h, const sqdk_space* sp, const void* coef, const void* v,
              void* out, int rows, int use_pc, int xcd, hipStream_t s, const int32_t* skip,
              int32_t* counter) {
  MatvecArgs<VT> args;
  args.keys = sp->keys;
  args.bstart = sp->bstart;
  args.fdir = sp->fdir;
  args.items = sp->items;
  args.pc_mask = sp->pc_mask;
  args.xmask = h->xmask;
  args.grp_off = h->grp_off;
  args.zmask = h->zmask;
  args.coef = reinterpret_cast<const typename CoefT<VT>::type*>(coef);
  args.class_hi = h->class_hi;
  args.class_off = h->class_off;
  args.v = reinterpret_cast<const VT*>(v);
  args.out = reinterpret_cast<VT*>(out);
  args.skip = skip;
  args.counter = counter;
  args.n_items = sp->n_items;
  args.m = sp->bucket_bits;
  args.d = sp->fine_bits;
  args.n_classes = h->n_classes;
  args.diag_t0 = h->diag_t0;
  args.diag_t1 = h->diag_t1;
  args.use_pc = use_pc;
  args.xcd_remap = xcd;
  const bool narrow = sp->bucket_bits <= 32;
  const size_t kt = narrow ? 4 : 8;
  const size_t lds = ((((size_t)(1 << sp->fine_bits) + 1) * 2 + 15) & ~(size_t)15) +
                     kt * (size_t)sp->max_bucket;
  if (lds > 160 * 1024)
    return fail(SQDK_ERANGE, "LDS tables of %zu bytes exceed 160 KiB (largest bucket %d keys)",
                lds, sp->max_bucket);
  return narrow ? launch_kt<VT, uint32_t>(args, rows, lds, s)
                : launch_kt<VT, uint64_t>(args, rows, lds, s);
}

}  // namespace

int matvec_impl(const sqdk_ham* ham, const sqdk_space* space, int32_t precision,
                const void* v, void* out, const sqdk_tune* tune, hipStream_t s,
                const int32_t* skip, int32_t* counter) {
  SQDK_TRY(check_space_public(space));
  SQDK_REQUIRE(ham != nullptr, "ham is NULL");
  SQDK_REQUIRE(v && out, "NULL vector");
  SQDK_REQUIRE(space->fdir && space->items && space->pc_mask, "space lacks fine directory or items");
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
  SQDK_REQUIRE(space->fine_bits >= 0 && space->fine_bits <= 14 &&
                   space->fine_bits <= space->bucket_bits,
               "fine_bits %d invalid", space->fine_bits);
  if (space->n_items == 0) return SQDK_OK;
  int rows = 8, use_pc = 1, xcd = 1;
  if (tune) {
    if (tune->rows_per_thread) rows = tune->rows_per_thread;
    use_pc = tune->use_pc_filter;
    xcd = tune->xcd_remap;
  }
  SQDK_REQUIRE(rows * kBlock == space->chunk_rows,
               "rows_per_thread %d does not match the space's chunk of %d rows", rows,
               space->chunk_rows);
  if (precision == SQDK_C128) {
    SQDK_REQUIRE(ham->n_terms == 0 || ham->coef, "ham has no complex128 coefficients");
    return launch_vt<double2>(ham, space, ham->coef, v, out, rows, use_pc, xcd, s, skip, counter);
  }
  if (precision == SQDK_C64) {
    SQDK_REQUIRE(ham->n_terms == 0 || ham->coef32, "ham has no complex64 coefficients");
    return launch_vt<float2>(ham, space, ham->coef32, v, out, rows, use_pc, xcd, s, skip, counter);
  }
  return fail(SQDK_EINVAL, "unknown precision %d", precision);
}

}  // namespace sqdk

extern "C" int sqdk_matvec(const sqdk_ham* ham, const sqdk_space* space, int32_t precision,
                           const void* v, void* out, const sqdk_tune* tune, void* stream) {
  return sqdk::matvec_impl(ham, space, precision, v, out, tune, sqdk::as_stream(stream));
}
