// State-vector simulation of rotation circuits (rqutils/svsim.py:96-122).
//
// A gate in symplectic form (X mask x, Z mask z, cos c, sin s) maps the state
// to c psi + i s Z^z X^x psi:
//   out[i] = c psi[i] + i s (-1)^{popcount(i & z)} psi[i ^ x],
// the update of do_svsim (the cos term is skipped when c == 0, as there).
//
// The reference makes one pass over the 2^n amplitudes per gate. Here the
// host splits the circuit into runs: consecutive gates whose X masks, above
// the lowest kMinLow qubits, touch at most k - kMinLow qubits. A run's tile
// qubits Q are those qubits plus the lowest qubits up to k in all; fixing
// every other qubit selects one tile of 2^k amplitudes, which a workgroup
// loads into LDS (contiguous runs of at least 2^kMinLow amplitudes, so loads
// stay coalesced), takes through every gate of the run, and writes back in
// place. Each run is then one read and one write of the state instead of one
// per gate. Within a full tile the amplitudes sit in registers (8 or 16 per
// thread) and most gates never touch LDS (svsim_reg_kernel); tiles of fewer
// qubits apply every gate in LDS (svsim_run_kernel). A gate with X on more
// high qubits than a tile holds takes one global pass of its own (pairs
// (i, i ^ x) per thread, in place).
#include "internal.hpp"

#include <vector>

namespace sqdk {

namespace {

constexpr int kSvThreads = 256;
// Contiguous low qubits in every tile: runs of 128 bytes (one cache line),
// 8 complex128 or 16 complex64 amplitudes.
template <typename VT> constexpr int kMinLow = sizeof(VT) == 16 ? 3 : 4;
constexpr int kMaxRunGates = 120;   // gates per launch (kernel argument space)

struct SvGate {
  uint32_t xl, zl;   // X and Z masks in tile-local bit positions
  uint64_t zo;       // Z bits outside the tile: a per-tile sign
  double c, s;
};

struct SvRun {
  uint64_t qmask;    // tile qubits
  int32_t n, k, low; // qubits, tile bits, contiguous low tile bits
  int32_t ngates;
  SvGate g[kMaxRunGates];
};

template <typename VT> struct SvReal;
template <> struct SvReal<double2> { using type = double; };
template <> struct SvReal<float2> { using type = float; };

// Bits of t deposited, in order, at the set positions of mask.
__device__ inline uint64_t deposit(uint64_t t, uint64_t mask) {
  uint64_t out = 0;
  while (mask) {
    const uint64_t low = mask & (~mask + 1);
    if (t & 1) out |= low;
    t >>= 1;
    mask &= mask - 1;
  }
  return out;
}

// out = c a + i s ph b (the cos term only when c != 0, as the reference).
template <typename VT>
__device__ inline VT rot(VT a, VT b, typename SvReal<VT>::type c, typename SvReal<VT>::type sp) {
  VT o;
  o.x = -sp * b.y;
  o.y = sp * b.x;
  if (c != 0) {
    o.x = c * a.x + o.x;
    o.y = c * a.y + o.y;
  }
  return o;
}

template <typename VT, int K>
__global__ void __launch_bounds__(kSvThreads)
svsim_run_kernel(VT* __restrict__ state, int64_t n_tiles, SvRun run) {
  using RT = typename SvReal<VT>::type;
  constexpr int kTile = 1 << K;
  __shared__ VT tile[kTile];
  __shared__ uint64_t hoff[1 << (K - kMinLow<VT>)];
  const int tid = threadIdx.x;
  const int low = run.low;
  const uint32_t lowmask = (1u << low) - 1u;
  const uint64_t nmask = run.n >= 64 ? ~0ull : ((1ull << run.n) - 1ull);
  const uint64_t hmask = run.qmask & ~(uint64_t)lowmask;   // tile qubits above the low block
  const uint64_t omask = nmask & ~run.qmask;              // qubits that pick the tile
  for (int h = tid; h < (1 << (run.k - low)); h += kSvThreads) hoff[h] = deposit((uint64_t)h, hmask);
  __syncthreads();
  const int size = 1 << run.k;
  for (int64_t t = blockIdx.x; t < n_tiles; t += gridDim.x) {
    const uint64_t base = deposit((uint64_t)t, omask);
    for (int e = tid; e < size; e += kSvThreads)
      tile[e] = state[base | hoff[e >> low] | (uint32_t)(e & lowmask)];
    __syncthreads();
    for (int gi = 0; gi < run.ngates; ++gi) {
      const SvGate G = run.g[gi];
      const RT c = (RT)G.c, s = (RT)G.s;
      const int tpar = __popcll(base & G.zo) & 1;
      if (G.xl == 0) {
        for (int e = tid; e < size; e += kSvThreads) {
          const RT sp = ((__popc(e & G.zl) ^ tpar) & 1) ? -s : s;
          tile[e] = rot(tile[e], tile[e], c, sp);
        }
      } else {
        const int b = __ffs(G.xl) - 1;   // lowest X bit: pairs (l, l ^ x) with bit b clear in l
        const uint32_t below = (1u << b) - 1u;
        for (int p = tid; p < size / 2; p += kSvThreads) {
          const uint32_t l = (((uint32_t)p & ~below) << 1) | ((uint32_t)p & below);
          const uint32_t m = l ^ G.xl;
          const VT a = tile[l], bb = tile[m];
          const RT sl = ((__popc(l & G.zl) ^ tpar) & 1) ? -s : s;
          const RT sm = ((__popc(m & G.zl) ^ tpar) & 1) ? -s : s;
          tile[l] = rot(a, bb, c, sl);
          tile[m] = rot(bb, a, c, sm);
        }
      }
      __syncthreads();
    }
    for (int e = tid; e < size; e += kSvThreads)
      state[base | hoff[e >> low] | (uint32_t)(e & lowmask)] = tile[e];
    __syncthreads();  // the next tile reuses the LDS buffer
  }
}

// One gate straight in global memory: pairs (i, i ^ x), or single amplitudes
// when x == 0.
template <typename VT>
__global__ void __launch_bounds__(kSvThreads)
svsim_gate_kernel(VT* __restrict__ state, int n, uint64_t x, uint64_t z, double cd, double sd) {
  using RT = typename SvReal<VT>::type;
  const RT c = (RT)cd, s = (RT)sd;
  const int64_t total = x ? (int64_t)1 << (n - 1) : (int64_t)1 << n;
  const int b = x ? __ffsll((long long)x) - 1 : 0;
  const uint64_t below = x ? ((1ull << b) - 1ull) : 0;
  for (int64_t p = blockIdx.x * (int64_t)kSvThreads + threadIdx.x; p < total;
       p += (int64_t)gridDim.x * kSvThreads) {
    if (x == 0) {
      const RT sp = (__popcll((uint64_t)p & z) & 1) ? -s : s;
      state[p] = rot(state[p], state[p], c, sp);
    } else {
      const uint64_t l = (((uint64_t)p & ~below) << 1) | ((uint64_t)p & below);
      const uint64_t m = l ^ x;
      const VT a = state[l], bb = state[m];
      const RT sl = (__popcll(l & z) & 1) ? -s : s;
      const RT sm = (__popcll(m & z) & 1) ? -s : s;
      state[l] = rot(a, bb, c, sl);
      state[m] = rot(bb, a, c, sm);
    }
  }
}

// ---- register-resident runs --------------------------------------------------
//
// The tile's 2^K amplitudes are spread over the 256 threads, 2^R each: a
// layout picks R local bit positions (the register bits); thread t holds the
// amplitudes whose other local bits spell t. Z-only gates and single-qubit X
// gates on a register bit then need no LDS traffic at all: the sign of an
// amplitude is parity(t & zt) ^ parity(e & zr) ^ the tile's outside parity,
// the second term uniform across the workgroup. The registers go through LDS
// only when a gate's X bit is not a register bit (a new layout) or a gate
// flips several tile bits (applied in LDS as in svsim_run_kernel).

// Op codes. Register gates: kRegDiag (Z-only) or kRegPair + j (X on register
// bit j), plus kRegNoCos when the gate's cos is 0 (the cos term is skipped, as
// in the reference); the code selects one straight-line variant per op.
enum SvOpCode : uint32_t {
  kRegPair = 0,      // + j, j < 4
  kRegDiag = 4,
  kRegNoCos = 8,     // added to kRegPair + j or kRegDiag
  kOpLds = 16,       // X on several tile bits: applied in LDS
  kOpLoad = 17,      // registers <- LDS in a new layout
  kOpStore = 18,     // registers -> LDS
};

// The first 32 bytes are all a register gate reads (one scalar load, issued
// one op ahead); the rest serves LDS gates and layout loads.
struct SvOp {
  double c, s;
  uint64_t zo;       // Z bits outside the tile
  uint32_t code;     // SvOpCode
  uint16_t zt;       // register ops: Z bits on the thread positions (thread-id coordinates)
  uint8_t zr;        // register ops: Z bits on the register positions (element coordinates)
  uint8_t pad0;
  uint32_t xl, zl;   // tile-local X and Z masks (LDS gates)
  uint8_t rpos[4];   // load: local position of each register bit, ascending
  uint32_t pad1;
};
static_assert(sizeof(SvOp) == 48, "op layout");

struct SvHead {
  double c, s;
  uint64_t zo;
  uint32_t code, zt, zr;
};

__device__ inline SvHead op_head(const SvOp& o) {
  SvHead h;
  h.c = o.c;
  h.s = o.s;
  h.zo = o.zo;
  h.code = o.code;
  h.zt = o.zt;
  h.zr = o.zr;
  return h;
}

constexpr int kMaxOps = 80;   // ops per launch (kernel argument space)

struct SvRunReg {
  uint64_t qmask;
  int32_t n, k, low, nops;
  SvOp op[kMaxOps];
};

// out = c a + i sp b, with the cos term present (C) or skipped.
template <bool C, typename VT>
__device__ inline VT rotc(VT a, VT b, typename SvReal<VT>::type c,
                          typename SvReal<VT>::type sp) {
  VT o;
  if constexpr (C) {
    o.x = __builtin_fma(-sp, b.y, c * a.x);
    o.y = __builtin_fma(sp, b.x, c * a.y);
  } else {
    o.x = -sp * b.y;
    o.y = sp * b.x;
  }
  return o;
}

// Z-only gate on every register: amplitude e takes sign parity(e & zr).
template <bool C, int R, typename VT>
__device__ inline void reg_diag(VT (&reg)[1 << R], typename SvReal<VT>::type c,
                                typename SvReal<VT>::type sp, typename SvReal<VT>::type sn,
                                uint32_t zr) {
#pragma unroll
  for (int e = 0; e < (1 << R); ++e) {
    const auto se = (__popc((uint32_t)e & zr) & 1) ? sn : sp;
    reg[e] = rotc<C>(reg[e], reg[e], c, se);
  }
}

// Single-qubit X on register bit B (a power of two): pairs (e, e | B).
template <bool C, int B, int R, typename VT>
__device__ inline void reg_pair(VT (&reg)[1 << R], typename SvReal<VT>::type c,
                                typename SvReal<VT>::type sp, typename SvReal<VT>::type sn,
                                uint32_t zr) {
#pragma unroll
  for (int e = 0; e < (1 << R); ++e) {
    if (e & B) continue;
    const int f = e | B;
    const auto se = (__popc((uint32_t)e & zr) & 1) ? sn : sp;
    const auto sf = (__popc((uint32_t)f & zr) & 1) ? sn : sp;
    const VT a = reg[e], b = reg[f];
    reg[e] = rotc<C>(a, b, c, se);
    reg[f] = rotc<C>(b, a, c, sf);
  }
}

template <typename VT, int K, int R>
__global__ void __launch_bounds__(kSvThreads)
svsim_reg_kernel(VT* __restrict__ state, int64_t n_tiles, SvRunReg run) {
  using RT = typename SvReal<VT>::type;
  constexpr int kTile = 1 << K;
  constexpr int E = 1 << R;
  static_assert(kTile == kSvThreads * E, "a tile is 256 threads x 2^R amplitudes");
  __shared__ VT tile[kTile];
  __shared__ uint64_t hoff[1 << (K - kMinLow<VT>)];
  const int tid = threadIdx.x;
  const int low = run.low;
  const int nops = run.nops;
  const uint32_t lowmask = (1u << low) - 1u;
  const uint64_t nmask = run.n >= 64 ? ~0ull : ((1ull << run.n) - 1ull);
  const uint64_t hmask = run.qmask & ~(uint64_t)lowmask;
  const uint64_t omask = nmask & ~run.qmask;
  for (int h = tid; h < (1 << (K - low)); h += kSvThreads) hoff[h] = deposit((uint64_t)h, hmask);
  __syncthreads();
  // Thread tid moves elements tid + 256 i of a tile. The next tile's
  // amplitudes are loaded into registers (pre) while this tile is processed.
  auto gidx = [&](uint64_t b, int i) {
    const int e = tid + i * kSvThreads;
    return b | hoff[e >> low] | (uint32_t)(e & lowmask);
  };
  // Tiles t, t + grid, ... of this block. The next tile's amplitudes are
  // loaded into registers (pre) while this tile is processed; at the end of a
  // tile its amplitudes go to registers (out), the next tile enters LDS, and
  // only then are the stores and the following loads issued, so no wait on a
  // load ever covers a store just issued. The loads past a block's last tile
  // read that tile again (in bounds, never used), so they need no condition.
  if ((int64_t)blockIdx.x >= n_tiles) return;
  RT pre_x[E], pre_y[E];   // plain reals: an array of VT is not kept in registers
  auto fetch = [&](uint64_t b) {
#pragma unroll
    for (int i = 0; i < E; ++i) {
      const VT a = state[gidx(b, i)];
      pre_x[i] = a.x;
      pre_y[i] = a.y;
    }
  };
  auto stage = [&]() {
#pragma unroll
    for (int i = 0; i < E; ++i) {
      VT a;
      a.x = pre_x[i];
      a.y = pre_y[i];
      tile[tid + i * kSvThreads] = a;
    }
  };
  int64_t t = blockIdx.x;
  uint64_t base = deposit((uint64_t)t, omask);
  fetch(base);
  stage();
  int64_t tn = t + gridDim.x < n_tiles ? t + gridDim.x : t;
  uint64_t nbase = deposit((uint64_t)tn, omask);
  fetch(nbase);
  for (;;) {
    // The ops come as segments: a layout load, register gates, a store; or
    // one LDS gate. The registers live within one segment only.
    int oi = 0;
    SvHead H = op_head(run.op[0]);
    while (oi < nops) {
      if (H.code == kOpLoad) {
        const SvOp& O = run.op[oi];
        __syncthreads();
        uint32_t rp[4] = {0, 0, 0, 0};
        uint32_t v = (uint32_t)tid;
#pragma unroll
        for (int i = 0; i < R; ++i) {  // insert a zero bit at each register position
          rp[i] = O.rpos[i];
          v = ((v >> rp[i]) << (rp[i] + 1)) | (v & ((1u << rp[i]) - 1u));
        }
        const uint32_t tb = v;       // thread part of the local index in this layout
        VT reg[E];
#pragma unroll
        for (int e = 0; e < E; ++e) {
          uint32_t d = 0;
#pragma unroll
          for (int i = 0; i < R; ++i) d |= (uint32_t)((e >> i) & 1) << rp[i];
          reg[e] = tile[tb | d];
        }
        ++oi;
        H = op_head(run.op[oi < nops ? oi : nops - 1]);
#pragma unroll 2
        while (oi < nops && H.code != kOpStore) {
          const SvHead g = H;
          ++oi;
          H = op_head(run.op[oi < nops ? oi : nops - 1]);   // the next op's header, early
          const RT c = (RT)g.c, s = (RT)g.s;
          // the thread's sign (thread bits and outside bits), then per element
          const int pt = (__popc((uint32_t)tid & g.zt) ^ __popcll(base & g.zo)) & 1;
          const RT sp = pt ? -s : s, sn = -sp;
          const uint32_t zr = g.zr;
          if constexpr (R > 3) {
            switch (g.code) {
              case kRegPair + 0: reg_pair<true, 1, R>(reg, c, sp, sn, zr); break;
              case kRegPair + 1: reg_pair<true, 2, R>(reg, c, sp, sn, zr); break;
              case kRegPair + 2: reg_pair<true, 4, R>(reg, c, sp, sn, zr); break;
              case kRegPair + 3: reg_pair<true, 8, R>(reg, c, sp, sn, zr); break;
              case kRegDiag: reg_diag<true, R>(reg, c, sp, sn, zr); break;
              case kRegNoCos + kRegPair + 0: reg_pair<false, 1, R>(reg, c, sp, sn, zr); break;
              case kRegNoCos + kRegPair + 1: reg_pair<false, 2, R>(reg, c, sp, sn, zr); break;
              case kRegNoCos + kRegPair + 2: reg_pair<false, 4, R>(reg, c, sp, sn, zr); break;
              case kRegNoCos + kRegPair + 3: reg_pair<false, 8, R>(reg, c, sp, sn, zr); break;
              default: reg_diag<false, R>(reg, c, sp, sn, zr); break;
            }
          } else {
            switch (g.code) {
              case kRegPair + 0: reg_pair<true, 1, R>(reg, c, sp, sn, zr); break;
              case kRegPair + 1: reg_pair<true, 2, R>(reg, c, sp, sn, zr); break;
              case kRegPair + 2: reg_pair<true, 4, R>(reg, c, sp, sn, zr); break;
              case kRegDiag: reg_diag<true, R>(reg, c, sp, sn, zr); break;
              case kRegNoCos + kRegPair + 0: reg_pair<false, 1, R>(reg, c, sp, sn, zr); break;
              case kRegNoCos + kRegPair + 1: reg_pair<false, 2, R>(reg, c, sp, sn, zr); break;
              case kRegNoCos + kRegPair + 2: reg_pair<false, 4, R>(reg, c, sp, sn, zr); break;
              default: reg_diag<false, R>(reg, c, sp, sn, zr); break;
            }
          }
        }
        // the store that closes the segment
#pragma unroll
        for (int e = 0; e < E; ++e) {
          uint32_t d = 0;
#pragma unroll
          for (int i = 0; i < R; ++i) d |= (uint32_t)((e >> i) & 1) << rp[i];
          tile[tb | d] = reg[e];
        }
        ++oi;
        H = op_head(run.op[oi < nops ? oi : nops - 1]);
      } else {
        // a gate with X on several tile bits, in LDS
        const SvOp& O = run.op[oi];
        const RT c = (RT)O.c, s = (RT)O.s;
        __syncthreads();
        const int tpar = __popcll(base & O.zo) & 1;
        const int b = __ffs(O.xl) - 1;
        const uint32_t below = (1u << b) - 1u;
        for (int p = tid; p < kTile / 2; p += kSvThreads) {
          const uint32_t l = (((uint32_t)p & ~below) << 1) | ((uint32_t)p & below);
          const uint32_t m = l ^ O.xl;
          const VT a = tile[l], bb = tile[m];
          const RT sl = ((__popc(l & O.zl) ^ tpar) & 1) ? -s : s;
          const RT sm = ((__popc(m & O.zl) ^ tpar) & 1) ? -s : s;
          tile[l] = rot(a, bb, c, sl);
          tile[m] = rot(bb, a, c, sm);
        }
        ++oi;
        H = op_head(run.op[oi < nops ? oi : nops - 1]);
      }
    }
    __syncthreads();
    RT out_x[E], out_y[E];
#pragma unroll
    for (int i = 0; i < E; ++i) {
      const VT a = tile[tid + i * kSvThreads];
      out_x[i] = a.x;
      out_y[i] = a.y;
    }
    __syncthreads();  // every thread has read this tile: LDS takes the next one
    stage();
#pragma unroll
    for (int i = 0; i < E; ++i) {
      VT a;
      a.x = out_x[i];
      a.y = out_y[i];
      state[gidx(base, i)] = a;
    }
    if (t + gridDim.x >= n_tiles) break;
    t += gridDim.x;
    base = nbase;
    tn = t + gridDim.x < n_tiles ? t + gridDim.x : t;
    nbase = deposit((uint64_t)tn, omask);
    fetch(nbase);
  }
}

// Position of each set bit of qmask, ascending: tile bit j sits at qubit pos[j].
int tile_positions(uint64_t qmask, int* pos) {
  int j = 0;
  for (int q = 0; q < 64; ++q)
    if ((qmask >> q) & 1ull) pos[j++] = q;
  return j;
}

uint32_t to_local(uint64_t m, const int* pos, int k) {
  uint32_t out = 0;
  for (int j = 0; j < k; ++j)
    if ((m >> pos[j]) & 1ull) out |= 1u << j;
  return out;
}

// Bits of m at the positions pos[0..cnt), packed in that order.
uint32_t extract_bits(uint32_t m, const int* pos, int cnt) {
  uint32_t out = 0;
  for (int i = 0; i < cnt; ++i)
    if ((m >> pos[i]) & 1u) out |= 1u << i;
  return out;
}

// Plans one run (tile qubits qmask, full tiles of 2^K) as register ops and
// launches it, one launch per kMaxOps ops; every launch is one pass over the
// state. A layout takes the X bit the gate needs, then the next gates' single
// X bits, then the highest tile bits (the low bits stay thread bits, which
// keeps the LDS accesses of a layout change contiguous across lanes).
template <typename VT, int K, int R>
int launch_reg_run(VT* state, int n, int low, uint64_t qmask, const int* pos,
                   const sqdk_gate* g, int64_t ng, int64_t n_tiles, unsigned grid, int& passes,
                   hipStream_t s) {
  static_assert(K - R == 8, "256 threads spell the thread bits");
  SvRunReg run;
  run.qmask = qmask;
  run.n = n;
  run.k = K;
  run.low = low;
  run.nops = 0;
  std::vector<uint32_t> xl(ng), zl(ng);
  for (int64_t t = 0; t < ng; ++t) {
    xl[t] = to_local(g[t].x, pos, K);
    zl[t] = to_local(g[t].z & qmask, pos, K);
  }
  int rset[4] = {0, 0, 0, 0};   // register bit positions of the current layout, ascending
  int tset[K];                  // the thread bit positions
  bool have = false;
  auto push = [&](const SvOp& op) { run.op[run.nops++] = op; };
  auto store = [&]() {
    SvOp op{};
    op.code = kOpStore;
    push(op);
    have = false;
  };
  auto flush = [&]() -> int {
    if (have) store();
    if (run.nops == 0) return SQDK_OK;
    svsim_reg_kernel<VT, K, R><<<grid, kSvThreads, 0, s>>>(state, n_tiles, run);
    SQDK_LAUNCHED("svsim_reg_kernel");
    ++passes;
    run.nops = 0;
    return SQDK_OK;
  };
  auto choose = [&](int b, int64_t t) {
    uint32_t set = b >= 0 ? 1u << b : 0u;
    for (int64_t u = t + 1; u < ng && u < t + 64 && __builtin_popcount(set) < R; ++u)
      if (__builtin_popcount(xl[u]) == 1) set |= xl[u];
    for (int p = K - 1; p >= 0 && __builtin_popcount(set) < R; --p) set |= 1u << p;
    int ri = 0, ti = 0;
    for (int p = 0; p < K; ++p) {
      if ((set >> p) & 1u) rset[ri++] = p;
      else tset[ti++] = p;
    }
    SvOp op{};
    op.code = kOpLoad;
    for (int i = 0; i < R; ++i) op.rpos[i] = (uint8_t)rset[i];
    push(op);
    have = true;
  };
  for (int64_t t = 0; t < ng; ++t) {
    if (run.nops + 4 > kMaxOps) SQDK_TRY(flush());   // room for store, load, op, store
    const int bits = __builtin_popcount(xl[t]);
    SvOp op{};
    op.xl = xl[t];
    op.zl = zl[t];
    op.zo = g[t].z & ~qmask;
    op.c = g[t].cos;
    op.s = g[t].sin;
    if (bits <= 1) {
      const int b = bits ? __builtin_ctz(xl[t]) : -1;
      bool in = have;
      if (in && b >= 0) {
        in = false;
        for (int i = 0; i < R; ++i) in = in || rset[i] == b;
      }
      if (!in) {
        if (have) store();
        choose(b, t);
      }
      uint32_t code = kRegDiag;
      for (int i = 0; i < R; ++i)
        if (bits && rset[i] == b) code = kRegPair + (uint32_t)i;
      // the cos term is skipped when cos is 0 in the kernel's precision
      if ((typename SvReal<VT>::type)g[t].cos == 0) code += kRegNoCos;
      op.code = code;
      op.zr = (uint8_t)extract_bits(zl[t], rset, R);
      op.zt = (uint16_t)extract_bits(zl[t], tset, K - R);
    } else {
      if (have) store();
      op.code = kOpLds;
    }
    push(op);
  }
  return flush();
}

template <typename VT, int K, int R>
int launch_runs(int n, VT* state, const sqdk_gate* gates, int64_t n_gates, int32_t* runs_out,
                hipStream_t s) {
  const int k = n < K ? n : K;
  const int min_low = k < kMinLow<VT> ? k : kMinLow<VT>;
  const uint64_t lowfix = (1ull << min_low) - 1ull;
  const uint64_t nmask = n >= 64 ? ~0ull : ((1ull << n) - 1ull);
  const int64_t n_tiles = (int64_t)1 << (n - k);
  const unsigned grid = (unsigned)std::min<int64_t>(n_tiles, 4096);
  int runs = 0;
  int64_t i = 0;
  while (i < n_gates) {
    const uint64_t hx0 = gates[i].x & ~lowfix;
    if (__builtin_popcountll(hx0) > k - min_low) {
      // X on more high qubits than a tile holds: one global pass.
      const int64_t total = gates[i].x ? (int64_t)1 << (n - 1) : (int64_t)1 << n;
      const unsigned g = (unsigned)std::min<int64_t>(ceil_div(total, kSvThreads), 8192);
      svsim_gate_kernel<VT><<<g, kSvThreads, 0, s>>>(state, n, gates[i].x, gates[i].z,
                                                      gates[i].cos, gates[i].sin);
      SQDK_LAUNCHED("svsim_gate_kernel");
      ++runs;
      ++i;
      continue;
    }
    // Greedy run: add gates while the union of their high X bits fits.
    uint64_t h = hx0;
    int64_t j = i + 1;
    while (j < n_gates && j - i < kMaxRunGates) {
      const uint64_t hn = h | (gates[j].x & ~lowfix);
      if (__builtin_popcountll(hn) > k - min_low) break;
      h = hn;
      ++j;
    }
    // Tile qubits: the run's high X bits, then the lowest other qubits.
    uint64_t q = h;
    for (int b = 0; b < n && __builtin_popcountll(q) < k; ++b) q |= 1ull << b;
    SvRun run;
    run.qmask = q & nmask;
    run.n = n;
    run.k = k;
    int low = 0;
    while (low < k && ((q >> low) & 1ull)) ++low;
    run.low = low;
    int pos[64];
    tile_positions(run.qmask, pos);
    if (k == K) {
      const int rc = launch_reg_run<VT, K, R>(state, n, low, run.qmask, pos, gates + i, j - i,
                                              n_tiles, grid, runs, s);
      if (rc != SQDK_OK) return rc;
      i = j;
      continue;
    }
    run.ngates = (int)(j - i);
    for (int64_t t = i; t < j; ++t) {
      SvGate& g = run.g[t - i];
      g.xl = to_local(gates[t].x, pos, k);
      g.zl = to_local(gates[t].z & run.qmask, pos, k);
      g.zo = gates[t].z & ~run.qmask;
      g.c = gates[t].cos;
      g.s = gates[t].sin;
    }
    svsim_run_kernel<VT, K><<<grid, kSvThreads, 0, s>>>(state, n_tiles, run);
    SQDK_LAUNCHED("svsim_run_kernel");
    ++runs;
    i = j;
  }
  if (runs_out) *runs_out = runs;
  return SQDK_OK;
}

}  // namespace

}  // namespace sqdk

extern "C" int sqdk_svsim(int32_t num_qubits, int32_t precision, void* state,
                          const sqdk_gate* gates, int64_t n_gates, int32_t* runs_out,
                          void* stream) {
  using namespace sqdk;
  SQDK_REQUIRE(num_qubits >= 1 && num_qubits <= 40, "num_qubits %d outside 1..40", num_qubits);
  SQDK_REQUIRE(state != nullptr, "NULL state");
  SQDK_REQUIRE(n_gates >= 0 && (n_gates == 0 || gates), "bad gate list");
  const uint64_t nmask = (1ull << num_qubits) - 1ull;
  for (int64_t i = 0; i < n_gates; ++i) {
    SQDK_REQUIRE(((gates[i].x | gates[i].z) & ~nmask) == 0,
                 "gate %lld acts on a qubit >= %d", (long long)i, num_qubits);
  }
  if (runs_out) *runs_out = 0;
  if (n_gates == 0) return SQDK_OK;
  if (precision == SQDK_C128)
    return launch_runs<double2, 11, 3>(num_qubits, static_cast<double2*>(state), gates, n_gates,
                                       runs_out, as_stream(stream));
  if (precision == SQDK_C64)
    return launch_runs<float2, 12, 4>(num_qubits, static_cast<float2*>(state), gates, n_gates,
                                      runs_out, as_stream(stream));
  return fail(SQDK_EINVAL, "unknown precision %d", precision);
}
