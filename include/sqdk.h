/*
 * sqdk.h -- C ABI of the MI355X SQD kernel library (librqsqd.so).
 *
 * The library implements the hot path of sample-based quantum diagonalization
 * (rqutils/sqd.py, rqutils/ground_locg.py): the matrix-free Pauli-sum matvec
 * over a sorted subspace of bitstrings and the single-vector LOBPCG that
 * drives it. Every entry point takes plain device pointers and sizes, returns
 * 0 on success or a negative status, and never throws; sqdk_last_error()
 * gives the message of the last failure on the calling thread. The library
 * allocates no device memory for data: callers pass buffers (sized by the
 * *_workspace queries) and a HIP stream (hipStream_t passed as void*).
 *
 * Layout conventions
 *   key     uint64: sum_q s_q 2^q of a bitstring (qubit q = bit q). Ascending
 *           integer order equals the reference's lexicographic order of
 *           bit-packed padded rows (rqutils/sqd.py:237, 424-426), so sorted
 *           order and source indices agree integer for integer.
 *   bucket  keys sharing key >> bucket_bits. bstart[b] is the index of the
 *           first key of bucket b; the aligned buckets are the unit of the
 *           matvec's LDS staging.
 *   group   Pauli terms sharing one X mask. Term k of group g lives at
 *           grp_off[g] <= k < grp_off[g+1] with Z mask zmask[k] and phased
 *           coefficient coef[2k] + i coef[2k+1] = c_k (-i)^{popcount(x&z)}
 *           (rqutils/paulis/symplectic.py:107-112).
 *   class   off-diagonal groups sharing x >> bucket_bits (the "high part"):
 *           for a row bucket a all of them read target bucket a ^ high.
 *   vectors complex, interleaved (re, im): double for complex128, float for
 *           complex64.
 */
#ifndef SQDK_H
#define SQDK_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SQDK_ABI_VERSION 20

enum sqdk_status {
  SQDK_OK = 0,
  SQDK_EINVAL = -1,   /* bad argument (maps to ValueError) */
  SQDK_ERANGE = -2,   /* size beyond a supported limit */
  SQDK_EHIP = -3,     /* HIP runtime error (maps to RuntimeError) */
  SQDK_ENOSPC = -4,   /* workspace or output buffer too small */
  SQDK_ECALLBACK = -5 /* an external matvec callback reported failure */
};

enum sqdk_precision { SQDK_C128 = 0, SQDK_C64 = 1 };

/* Sorted subspace S with its bucket directories (device pointers). Replaces
 * the uniquified state array states_u of rqutils/sqd.py:316 (rows past the
 * unique count are not stored: n is the unique count). */
typedef struct sqdk_space {
  const uint64_t* keys;    /* [n] strictly ascending */
  const uint32_t* bstart;  /* [2^(num_qubits - bucket_bits) + 1] */
  const uint16_t* fdir;    /* [2^(num_qubits - bucket_bits) << fine_bits]:
                              keys of bucket b below sub-bucket e */
  const uint32_t* items;   /* [n_items] first row of each row chunk */
  const uint64_t* pc_mask; /* [2]: bit w set iff some key has popcount w (host pruning) */
  int64_t n;               /* 1 <= n < 2^31 */
  int32_t num_qubits;      /* key width, 1..64 */
  int32_t bucket_bits;     /* m: bucket = key >> m; num_qubits - m <= 26 */
  int32_t fine_bits;       /* d <= min(m, 15) */
  int32_t chunk_rows;      /* rows per item; items never cross a bucket */
  int32_t n_items;
  int32_t max_bucket;      /* keys in the largest bucket (<= 16384) */
  const uint32_t* fbits;   /* NULL, or with fine_bits == bucket_bits - 5 the occupancy
                              bitmap of every 32-value sub-bucket (same index as fdir):
                              the matvec then looks targets up in O(1) instead of
                              searching (sqdk_space_bits) */
} sqdk_space;

/* Pauli sum grouped by X mask and ordered by class (device pointers).
 * Replaces the packed arrays of PauliSumXZ (rqutils/paulis/symplectic.py:
 * 42-49): x u8[J,Bp] -> xmask, z u8[J,Kmax,Bp] -> zmask + grp_off (ragged
 * instead of zero-padded), c [J,Kmax] -> coef / coef32. */
typedef struct sqdk_ham {
  const uint64_t* xmask;    /* [J] off-diagonal groups in class order */
  const int32_t* grp_off;   /* [J+1] term ranges of those groups */
  const uint64_t* zmask;    /* [T] */
  const double* coef;       /* [2T] complex128 phased coefficients */
  const float* coef32;      /* [2T] the same rounded to complex64 (or NULL) */
  const uint32_t* class_hi; /* [C] x >> bucket_bits of each class */
  const int32_t* class_off; /* [C+1] group ranges of the classes */
  int32_t n_groups;         /* J (off-diagonal groups only) */
  int32_t n_terms;          /* T (all terms, diagonal ones included) */
  int32_t n_classes;        /* C */
  int32_t diag_t0, diag_t1; /* term range of the X = 0 group (empty if none) */
  int32_t num_qubits;
  int32_t bucket_bits;      /* must equal the space's */
  int32_t reserved;
} sqdk_ham;

/* One instance of a batch (see sqdk_batch): where its tables start in the
 * concatenated arrays. All row, group and term indices stored in the tables
 * are global (they already include the instance's offsets). */
typedef struct sqdk_inst {
  int64_t bstart_base;      /* first entry of its bucket directory in bstart */
  int64_t fdir_base;        /* first entry of its fine directory in fdir */
  int32_t class_base;       /* first of its classes in class_hi / class_off */
  int32_t n_classes;
  int32_t diag_t0, diag_t1; /* term range of its X = 0 group */
  int32_t bucket_bits;
  int32_t fine_bits;
  int64_t fbits_base;       /* first word of its dense-lookup bitmaps in the batch's fbits
                               (fine_bits == bucket_bits - 5), or -1: search lookup */
} sqdk_inst;

/* Many independent instances (each its own Hamiltonian and subspace) laid
 * end to end, for one persistent matvec launch and a batched LOBPCG. Rows of
 * instance i are row_base[i] <= r < row_base[i+1]; vectors are the
 * concatenation of the instances' vectors. No counterpart in the reference:
 * it solves one instance per call. */
typedef struct sqdk_batch {
  const uint64_t* keys;      /* [n] keys, strictly ascending within each instance */
  const uint32_t* bstart;    /* concatenated bucket directories (global rows) */
  const uint16_t* fdir;      /* concatenated fine directories */
  const uint32_t* items;     /* [n_items] first (global) row of each chunk */
  const uint16_t* item_inst; /* [n_items] instance of each chunk */
  const int32_t* item_base;  /* [n_inst+1] chunks of instance i: item_base[i].. */
  const int64_t* row_base;   /* [n_inst+1] rows of instance i: row_base[i].. */
  const uint64_t* xmask;     /* concatenated group tables (global indices) */
  const int32_t* grp_off;
  const uint64_t* zmask;
  const double* coef;
  const float* coef32;
  const uint32_t* class_hi;
  const int32_t* class_off;
  const sqdk_inst* inst;     /* [n_inst] descriptors (device memory) */
  int64_t n;                 /* total rows */
  int32_t n_inst;
  int32_t n_items;
  int32_t chunk_rows;        /* rows per chunk: 256 x rows_per_thread */
  int32_t max_bucket;        /* largest bucket over the search-lookup instances (>= 1) */
  int32_t max_fine_bits;     /* largest fine_bits over the search-lookup instances */
  int32_t max_bucket_bits;   /* <= 32 selects 32-bit LDS keys */
  const uint32_t* fbits;     /* concatenated dense-lookup bitmaps, or NULL: every instance
                                searches. With it, one matvec is two launches over the
                                items, one per lookup, each skipping the other's instances */
  int32_t dense_fine_bits;   /* largest fine_bits over the dense-lookup instances */
  int32_t reserved;
  const struct sqdk_cache* cache; /* or NULL: the batched solve's products read this hit
                                     table (sqdk_batch_cache_*) instead of joining */
} sqdk_batch;

/* Launch tuning of the bucket-join matvec (0 = library default). */
typedef struct sqdk_tune {
  int32_t rows_per_thread; /* 4, 8 or 16; must match chunk_rows / 256 */
  int32_t group_split;     /* workgroups sharing a row chunk, groups dealt round
                              robin (<= 1: one); needs sqdk_matvec_workspace bytes */
  int32_t xcd_remap;       /* map consecutive items onto one XCD */
  int32_t stage_v;         /* stage each target bucket's slice of v in LDS */
  int32_t compact;         /* apply hits through a per-wave LDS queue (4 or 8 rows
                              per thread): one lane per hit instead of one per row */
} sqdk_tune;

/* ---- library ----------------------------------------------------------- */

/* No reference counterpart: ABI handshake for the Python loader. */
int sqdk_abi_version(void);

/* Thread-local message of the last failing call (empty string if none). */
const char* sqdk_last_error(void);

/* ---- subspace ---------------------------------------------------------- */

/* Check keys and summarise them: err_out[0] (device int32) receives the
 * number of out-of-order (not strictly ascending) or too-wide keys, pc_mask
 * (device, 2 words) the popcounts present. No reference counterpart: the
 * reference's inputs are uniquified by sorting (rqutils/sqd.py:414-445). */
int sqdk_space_scan(const uint64_t* keys, int64_t n, int32_t num_qubits,
                    uint64_t* pc_mask, int32_t* err_out, void* stream);

/* dir[p] = number of keys < (p << shift) for 0 <= p <= 2^(num_qubits-shift)
 * (uint32). Used for the bucket directory and to choose bucket_bits. */
int sqdk_space_dir(const uint64_t* keys, int64_t n, int32_t num_qubits,
                   int32_t shift, uint32_t* dir, void* stream);

/* Fine directory: fdir[(b << fine_bits) + e] = number of keys of bucket b
 * whose sub-bucket ((key >> (bucket_bits - fine_bits)) & (2^fine_bits - 1))
 * is below e. Buckets must hold fewer than 65536 keys. */
int sqdk_space_fine(const uint64_t* keys, int64_t n, int32_t num_qubits,
                    int32_t bucket_bits, int32_t fine_bits,
                    const uint32_t* bstart, uint16_t* fdir, void* stream);

/* ---- intake of sampled bitstrings (rqutils/sqd.py:237, 246, 414-445) ---- */

/* keys[r] = sum_c (states[r*row_stride + c] != 0) << (num_qubits - 1 - c): column
 * c of a row is qubit num_qubits-1-c, so ascending keys are the reference's
 * order of packed rows (sqd.py:237). */
int sqdk_intake_pack(const uint8_t* states, int64_t n_rows, int32_t num_qubits,
                     int64_t row_stride, uint64_t* keys, void* stream);

/* Sorted distinct keys: out[0..*n_unique) (out holds n entries; *n_unique is
 * written on the device, -1 if some key is not below 2^key_bits). Replaces the
 * lexsort + unique of uniquify_states (sqd.py:414-445), whose fill rows are
 * not materialised. Up to 36-bit keys, when 2^key_bits / 8 bytes are at most
 * 64 MiB or 32 bytes per key, an occupancy bitmap gives the result without a
 * sort; otherwise a stable LSD radix sort over the low key_bits bits and a
 * compaction. Workspace: sqdk_intake_unique_workspace (it depends on both n
 * and key_bits). */
int sqdk_intake_unique_workspace(int64_t n, int32_t key_bits, size_t* bytes);
int sqdk_intake_unique(const uint64_t* keys, int64_t n, int32_t key_bits, uint64_t* out,
                       int64_t* n_unique, void* work, size_t work_bytes, void* stream);

/* states[r*num_qubits + c] = bit (num_qubits - 1 - c) of keys[r]: the
 * basis_states the reference returns (sqd.py:246). */
int sqdk_intake_unpack(const uint64_t* keys, int64_t n, int32_t num_qubits, uint8_t* states,
                       void* stream);

/* Occupancy bitmaps for the dense lookup: with fine_bits == bucket_bits - 5,
 * fbits[e] has bit j set iff key (e << 5) | j is in the subspace, for every
 * fine-directory entry e (2^(num_qubits - 5) entries; needs fdir built first).
 * No counterpart in the reference (its lookup is a sort-join,
 * rqutils/sqd.py:448-503). */
int sqdk_space_bits(const uint64_t* keys, int64_t n, int32_t num_qubits, int32_t bucket_bits,
                    int32_t fine_bits, const uint32_t* bstart, const uint16_t* fdir,
                    uint32_t* fbits, void* stream);

/* Row chunks of at most chunk_rows rows that never cross a bucket boundary:
 * items[k] = first row of chunk k; the count goes to n_items_out (device
 * int32). Fails with SQDK_ENOSPC (count still written) if max_items is short. */
int sqdk_space_items(const uint32_t* bstart, int64_t n_buckets,
                     int32_t chunk_rows, uint32_t* items, int64_t max_items,
                     int32_t* n_items_out, void* stream);

/* xsource[i] = j with keys[j] == keys[i] ^ xmask, else -1 (int32).
 * Replaces get_xsource (rqutils/sqd.py:448-503); bit-exact. */
int sqdk_xsource(const sqdk_space* space, uint64_t xmask, int32_t* out,
                 void* stream);

/* Coefficient vector of one X group: out[i] = sum_k coef_k
 * (-1)^{popcount(zmask_k & key_i)} over terms t0 <= k < t1, complex128
 * (2n doubles). Replaces get_diagonal (rqutils/sqd.py:554-575); zero
 * coefficients are skipped, not a stop (the reference stops at them). */
int sqdk_diagonal(const uint64_t* zmask, const double* coef, int32_t t0,
                  int32_t t1, const uint64_t* keys, int64_t n, double* out,
                  void* stream);

/* ---- matrix-free matvec ------------------------------------------------ */

/* Workspace bytes of sqdk_matvec for this tuning (0 unless group_split > 1). */
int sqdk_matvec_workspace(const sqdk_space* space, const sqdk_tune* tune,
                          int32_t precision, size_t* bytes);

/* out = H v over the subspace: per row, the X = 0 group's coefficient times
 * v[row] plus, per off-diagonal group, one XOR and one search of the target
 * in the LDS-staged target bucket. Replaces apply_h / apply_h_x_cached and
 * siblings (rqutils/sqd.py:590-684). precision SQDK_C128: double storage and
 * FP64 sums; SQDK_C64: float storage, FP32 two-sum accumulation. */
int sqdk_matvec(const sqdk_ham* ham, const sqdk_space* space, int32_t precision,
                const void* v, void* out, const sqdk_tune* tune, void* work,
                size_t work_bytes, void* stream);

/* Workspace bytes of sqdk_batch_matvec for this tuning. */
int sqdk_batch_matvec_workspace(const sqdk_batch* batch, const sqdk_tune* tune,
                                int32_t precision, size_t* bytes);

/* out = H_i v_i for every instance i of the batch, in one launch. */
int sqdk_batch_matvec(const sqdk_batch* batch, int32_t precision, const void* v,
                      void* out, const sqdk_tune* tune, void* work, size_t work_bytes,
                      void* stream);

/* ---- vector kernels (LOBPCG building blocks) --------------------------- */

/* out[k] = sum_i conj(a_k[i]) b_k[i] for k < npairs (<= 16), accumulated in
 * double in a fixed order; results (complex double, 2*npairs) go to device
 * memory. Replaces the jnp.sum(conj(a)*b) reductions of
 * rqutils/ground_locg.py:214-421. work: sqdk_dots_workspace bytes. */
int sqdk_dots(int64_t n, int32_t precision, int32_t npairs,
              const void* const* a, const void* const* b, double* out,
              void* work, size_t work_bytes, void* stream);
int sqdk_dots_workspace(int64_t n, int32_t npairs, size_t* bytes);

/* sqdk_dots with no workspace argument and no allocation of any kind: the
 * per-block partial sums go to a file-scope device array of the library
 * (128 KiB on every GPU it runs on, held until the process ends), so npairs
 * is limited to the 16 pairs that array holds (larger requests fail with
 * SQDK_EINVAL). No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc). The stream's device must be the
 * current device; calls that overlap on one GPU share the array and race. */
int sqdk_dots_noalloc(int64_t n, int32_t precision, int32_t npairs, const void* const* a,
                      const void* const* b, double* out, void* stream);

/* Convert between complex128 and complex64 vectors (n elements). */
int sqdk_convert(int64_t n, int32_t from_precision, const void* in,
                 int32_t to_precision, void* out, void* stream);

/* ---- cached hit table (the sparse form of the reference's source cache) ----
 *
 * The reference can cache each group's source indices, 4 bytes per (row,
 * group) (cache_level[0] == 1, rqutils/sqd.py:353-371, 448-503). Here only the
 * hits are kept: for every row its (source row, group coefficient C_g(s_i))
 * pairs, found once by the bucket join, plus the X = 0 group as a diagonal.
 * Each window of 256 consecutive rows of a join chunk is sorted by hit count,
 * most first (equal counts, and the rows without hits, keep their order; slots
 * past the chunk's rows come last), and every 64 sorted slots form a slice; a
 * slice stores its
 * entries column-major (entry k of its lane l at base + 64 k + l) up to its
 * largest hit count, shorter rows padded with column 0xFFFFFFFF and
 * coefficient 0; row_len says where each row's hits end. Building:
 * sqdk_cache_slices (host count of slices), sqdk_cache_count (device: row
 * counts, the sort, the entry offsets, the last one the entry count),
 * allocation by the caller, sqdk_cache_fill. */
typedef struct sqdk_cache {
  int64_t n;                    /* rows (the subspace size) */
  int64_t n_slices;
  const uint32_t* slot_row;     /* [64 n_slices]: the row of each slot, 0xFFFFFFFF = none */
  const uint64_t* slice_base;   /* [n_slices + 1]: first entry; the last = entries */
  const uint32_t* col;          /* [entries] source row, 0xFFFFFFFF = padding */
  const void* coef;             /* [entries] complex coefficient (cache precision) */
  const void* diag;             /* [n] complex X = 0 coefficient of each row */
  const uint32_t* row_len;      /* [n] hits of each row (its entries before the padding) */
  int32_t precision;            /* SQDK_C128 or SQDK_C64 */
  int32_t slices_per_wave;      /* product: 1 (or 0) where rows have many hits, 4 where few
                                   (the builder's choice: 4 at most 8 entries per row) */
} sqdk_cache;

/* Slices of the subspace: n_items x chunk_rows / 64 (chunks of 1024, 2048 or
 * 4096 rows). */
int sqdk_cache_slices(const sqdk_space* space, int64_t* n_slices);
/* Bytes of the optional build workspace (device): n_items x (ceil(n_classes / 8)
 * + 1) words of class masks, then 2 x 4 bytes per row. The count records in it,
 * per chunk and class, which of the chunk's four waves had a hit, which waves
 * hold a row of more than two hits, and every row's first two source rows; the
 * fill then takes the entries of rows from waves without such a row from that
 * record and joins only the (wave, class) pairs with hits of the others. */
int sqdk_cache_workspace(const sqdk_ham* ham, const sqdk_space* space, size_t* bytes);
/* slot_row [64 n_slices], slice_base [n_slices + 1] and row_len [n] (device);
 * work: the build workspace (written here, read by sqdk_cache_fill) or NULL. */
int sqdk_cache_count(const sqdk_ham* ham, const sqdk_space* space, uint32_t* slot_row,
                     uint64_t* slice_base, uint32_t* row_len, void* work, size_t work_bytes,
                     void* stream);
/* col / coef [slice_base[n_slices]], diag [n] (device; coef and diag complex of
 * precision) at the slots of sqdk_cache_count; row_len [n] is rewritten.
 * work: the workspace the same table's sqdk_cache_count wrote, or NULL (the
 * fill then joins every class); the entries are the same either way. */
int sqdk_cache_fill(const sqdk_ham* ham, const sqdk_space* space, int32_t precision,
                    const uint64_t* slice_base, const uint32_t* slot_row, uint32_t* col,
                    void* coef, void* diag, uint32_t* row_len, const void* work,
                    size_t work_bytes, void* stream);
/* out = H v through the table (v, out: n complex of the cache's precision);
 * sums in FP64, each row's in group order; a row reads its own hits only. */
int sqdk_cache_matvec(const sqdk_cache* cache, const void* v, void* out, void* stream);

/* The same table for a whole batch (rows and slices global, one table): slices
 * = n_items x chunk_rows / 64; count and fill as above. The product is
 * sqdk_cache_matvec; sqdk_batch.cache hands it to the batched solve. The build
 * workspace (or NULL) is sized for max_classes (the largest n_classes of the
 * instances), the same in workspace, count and fill; classes of an instance past
 * max_classes are never skipped. */
int sqdk_batch_cache_slices(const sqdk_batch* batch, int64_t* n_slices);
int sqdk_batch_cache_workspace(const sqdk_batch* batch, int32_t max_classes, size_t* bytes);
int sqdk_batch_cache_count(const sqdk_batch* batch, uint32_t* slot_row, uint64_t* slice_base,
                           uint32_t* row_len, int32_t max_classes, void* work,
                           size_t work_bytes, void* stream);
int sqdk_batch_cache_fill(const sqdk_batch* batch, int32_t precision, const uint64_t* slice_base,
                          const uint32_t* slot_row, uint32_t* col, void* coef, void* diag,
                          uint32_t* row_len, int32_t max_classes, const void* work,
                          size_t work_bytes, void* stream);

/* ---- operator functions on the reference's own arrays ------------------- *
 * For callers that hold the reference's packed arrays. A packed padded row of
 * row_bytes <= 8 bytes (rqutils/sqd.py:237; a packed X or Z signature likewise,
 * rqutils/paulis/symplectic.py:117-124) is read as one big-endian integer, its
 * "row key": row-key order is the reference's lexicographic row order (0xFF
 * fill rows last), XOR and AND act on row keys exactly as on the packed bytes.
 */

/* keys[i] = big-endian integer of rows[i*row_bytes .. +row_bytes). */
int sqdk_rows_to_keys(const uint8_t* rows, int64_t n, int32_t row_bytes, uint64_t* keys,
                      void* stream);
/* rows[i] = the big-endian bytes of keys[i] for i < n_keys, 0xFF bytes for
 * n_keys <= i < n_rows: the fill rows of uniquify_states (sqd.py:441-445). */
int sqdk_keys_to_rows(const uint64_t* keys, int64_t n_keys, int64_t n_rows, int32_t row_bytes,
                      uint8_t* rows, void* stream);

enum sqdk_coef_form {
  SQDK_COEF_TERMS = 0, /* C_g from the group's Z signatures and coefficients (get_diagonal) */
  SQDK_COEF_DIAG = 1,  /* C_g given per row (cache_level[1] == 2: diagonals) */
  SQDK_COEF_SIGNS = 2  /* C_g from cached sign bits and the coefficients (cache_level[1] == 1) */
};

/* The reference's group arrays for one product (device pointers). */
typedef struct sqdk_groups {
  const uint64_t* keys;    /* [n] row keys of the states, sorted, fill rows (pad bit set) last */
  int64_t n;               /* rows, fill rows included (< 2^31) */
  int64_t n_real;          /* rows before the first fill row */
  int32_t n_groups;        /* J */
  int32_t coef_form;       /* sqdk_coef_form */
  const uint64_t* xmask;   /* [J] X signatures as row keys: used when xsource is NULL */
  const int32_t* xsource;  /* [J][ld] cached source indices (get_xsource) or NULL: each
                              (row, group) target is found by a binary search of keys */
  int64_t ld;              /* row stride of xsource, diag and signs (>= n) */
  const uint64_t* zmask;   /* TERMS: [J][kmax] Z signatures as row keys */
  const double* coef;      /* TERMS, SIGNS: [J][kmax] complex128 phased coefficients,
                              zero-padded; zero terms are skipped (the reference stops at
                              the first one, rqutils/sqd.py:562-564) */
  int32_t kmax;
  int32_t sign_bytes;      /* SIGNS: bytes per (group, row), >= kmax / 8 */
  const double* diag;      /* DIAG: [J][ld] complex128 */
  const uint8_t* signs;    /* SIGNS: [J][ld][sign_bytes], bit 7 - k % 8 of byte k / 8 */
} sqdk_groups;

/* out[i] = sum_g C_g(s_i) v[source_g(i)] in group order with FP64 sums; fill
 * rows get 0. Replaces apply_h (xsource NULL, TERMS; rqutils/sqd.py:590-605),
 * apply_h_s_cached (NULL, SIGNS; :608-622), apply_h_z_cached (NULL, DIAG;
 * :625-637), apply_h_x_cached (given, TERMS; :640-654), apply_h_xs_cached
 * (given, SIGNS; :657-670) and apply_h_xz_cached (given, DIAG; :673-684).
 * v and out: n complex elements of `precision`. */
int sqdk_groups_matvec(const sqdk_groups* g, int32_t precision, const void* v, void* out,
                       void* stream);

/* out[i*sign_bytes + k/8] bit 7 - k%8 = parity of zmask[k] & keys[i] for
 * k < n_terms (one group's terms). Replaces get_diag_signs (sqd.py:506-528). */
int sqdk_diag_signs(const uint64_t* zmask, int32_t n_terms, const uint64_t* keys, int64_t n,
                    int32_t sign_bytes, uint8_t* out, void* stream);

/* out[i] (complex128) = sum_k coef_k (1 - 2 bit_k(i)) over one group's terms.
 * Replaces compute_diagonal (sqd.py:531-551) with the bit of term k read at
 * 7 - k % 8 (the reference reads 7 - (k & 255): +1 for every term from the
 * ninth on). */
int sqdk_compute_diagonal(const uint8_t* signs, int64_t n, int32_t sign_bytes, const double* coef,
                          int32_t n_terms, double* out, void* stream);

/* ---- single-vector LOBPCG ---------------------------------------------- */

/* External matvec: write A*in into out (both n complex elements of the
 * solver's precision) on the given stream; return 0 on success. */
typedef int (*sqdk_matvec_fn)(void* ctx, const void* in, void* out, void* stream);

enum sqdk_locg_mode {
  SQDK_LOCG_REFERENCE = 0, /* 4 matvecs/iteration, structure of ground_locg.py:288-341 */
  SQDK_LOCG_FUSED = 1      /* 1 matvec/iteration, implicit products, Gram-based RR */
};

enum sqdk_locg_criterion {
  SQDK_CRIT_REFERENCE = 0, /* |r| < tol * 10 * n_for_tol * (|Ax| - theta), ground_locg.py:324-331 */
  SQDK_CRIT_RELATIVE = 1   /* |A x - theta x| <= tol * |theta| */
};

/* Preconditioner of the FUSED mode (the reference's iteration has none). */
enum sqdk_locg_precond {
  SQDK_PRECOND_NONE = 0,   /* search direction = residual R */
  SQDK_PRECOND_DIAG = 1    /* R / max(d - min(theta, min d), floor), d = diag(H) from the
                              X = 0 terms, floor = max(0.02 (mean d - min d), 1e-3 |theta|):
                              Jacobi with the shift kept at or below min d; needs the
                              built-in operator */
};

typedef struct sqdk_locg_opts {
  int32_t mode;            /* sqdk_locg_mode */
  int32_t precision;       /* SQDK_C128 or SQDK_C64 */
  int32_t criterion;       /* sqdk_locg_criterion */
  int32_t maxiter;
  double tol;
  int64_t n_for_tol;       /* vector length used by SQDK_CRIT_REFERENCE */
  int32_t check_every;     /* host convergence poll period in iterations (>=1) */
  int32_t refresh_every;   /* FUSED: recompute A x and A p explicitly every k iterations */
  int32_t use_graph;       /* FUSED with the built-in operator: capture one refresh
                              period (refresh_every iterations and the refresh) as a
                              HIP graph, replayed on a private stream after the
                              caller's stream is synchronised; the host polls once
                              per period. The batched solve captures refresh_every
                              iterations the same way */
  int32_t precond;         /* sqdk_locg_precond (FUSED only) */
  int32_t reserved;
  double* debug;           /* REFERENCE only, or NULL: device buffer of (maxiter + 2) x 30
                              doubles that receives one diagnostics record after the
                              first step, after the 2x2 step and after every loop
                              iteration (ground_locg.py:211-233, 264-286, 359-364):
                              theta, rho = Re <x, A x>, reltol = (|Ax| - theta) 10 N,
                              converged (0/1), |r|, 0, kappa[3] (complex), then S^H A S
                              of (x, y, r) (3x3 complex, row-major), at the cost of
                              three more products per record. As in the reference's
                              diagnostics mode every one of maxiter iterations runs;
                              the convergence test is recorded, not acted on */
  void* debug_vectors;     /* or NULL (needs debug): (maxiter + 2) x 3 x n elements of the
                              solver precision receiving x, y, r at every record */
} sqdk_locg_opts;

typedef struct sqdk_locg_result {
  double eigval;           /* Rayleigh-Ritz value theta of the returned vector */
  double resid_norm;       /* |r| of the last iterate as tracked by the solver */
  double ax_norm;          /* |A x| of the last iterate */
  int32_t niter;           /* loop iterations, counted as ground_locg.py:366-372 */
  int32_t converged;
  int32_t matvecs;
  int32_t status;          /* 0; 2: Rayleigh-Ritz breakdown (Cholesky pivot below the
                              floor with P already dropped); 3: NaN or Inf in the
                              sums, norms or small eigenproblems (non-finite operator
                              or vectors; converged is then 0); 4: the start is an
                              exact eigenvector and n > 1 (its residual is 0: an
                              invariant subspace such as a one-hot on a state with
                              no off-diagonal neighbour, so eigval need not be the
                              lowest eigenvalue; the reference divides 0 by 0 here,
                              ground_locg.py:274-286, 424-435) */
} sqdk_locg_result;

/* The operator: the built-in Pauli-sum matvec (ham + space + tune) when fn is
 * NULL, else the external callback. With cache set (and fn NULL) the products
 * use the cached hit table instead of the bucket join (its precision must be
 * the solve's). */
typedef struct sqdk_op {
  const sqdk_ham* ham;
  const sqdk_space* space;
  const sqdk_tune* tune;
  sqdk_matvec_fn fn;
  void* ctx;
  const struct sqdk_cache* cache;
} sqdk_op;

/* Device workspace (bytes) for a solve of n elements with this operator. */
int sqdk_locg_workspace(const sqdk_op* op, int64_t n, const sqdk_locg_opts* opts,
                        size_t* bytes);

/* Single-vector LOBPCG for the lowest eigenpair of a Hermitian operator.
 * Replaces ground_locg / _ground_locg_callable (rqutils/ground_locg.py:128-377).
 * x (n complex, solver precision, device) holds the start vector on entry and
 * the normalised eigenvector on return. */
int sqdk_locg_solve(const sqdk_op* op, int64_t n, void* x,
                    const sqdk_locg_opts* opts, void* work, size_t work_bytes,
                    sqdk_locg_result* result, void* stream);

/* Device workspace (bytes) of a batched solve. */
int sqdk_batch_locg_workspace(const sqdk_batch* batch, const sqdk_tune* tune,
                              const sqdk_locg_opts* opts, size_t* bytes);

/* Fused single-vector LOBPCG on every instance of the batch at once (mode
 * must be SQDK_LOCG_FUSED). x holds the concatenated start vectors on entry
 * and the normalised eigenvectors on return; results[i] (host array of
 * n_inst entries) receives instance i's eigenvalue, iterations and residual.
 * opts->use_graph captures refresh_every iterations as one HIP graph (on a
 * private stream, after synchronising the given one). No counterpart in the
 * reference, which solves one instance per ground_locg call. */
int sqdk_batch_locg_solve(const sqdk_batch* batch, const sqdk_tune* tune, void* x,
                          const sqdk_locg_opts* opts, void* work, size_t work_bytes,
                          sqdk_locg_result* results, void* stream);

/* Lowest eigenpair of one Hermitian 3x3 (Kopp 2008, as ground_locg.py:438-471)
 * or 2x2 (ground_locg.py:424-435) matrix, evaluated by the device code path:
 * mat (device) is 2*dim*dim doubles (row-major complex); out (device) receives
 * the eigenvalue then 2*dim doubles of eigenvector. dim = 2 or 3. */
int sqdk_small_eigenpair(int32_t dim, const double* mat, double* out,
                         void* stream);

/* ---- start-vector network (BF16 MLP on the matrix cores) --------------- *
 * No counterpart in the reference (it starts LOBPCG from a one-hot vector,
 * rqutils/sqd.py:373-398). bf16 values are passed as uint16 bit patterns. */

/* C = op(A) op(B) (+ bias[N]) with bf16 A, B and FP32 accumulation on MFMA.
 * op(A)(i,k) = trans_a ? A[k*lda + i] : A[i*lda + k]; op(B)(k,j) likewise
 * with ldb. ksplit > 1 splits K over that many partial tiles, reduced in a
 * fixed order (workspace: sqdk_gemm_bf16_workspace). */
int sqdk_gemm_bf16_workspace(int64_t M, int64_t N, int32_t ksplit, size_t* bytes);
int sqdk_gemm_bf16(int64_t M, int64_t N, int64_t K, const uint16_t* A, int64_t lda,
                   int32_t trans_a, const uint16_t* B, int64_t ldb, int32_t trans_b,
                   float* C, int64_t ldc, const float* bias, int32_t ksplit, float* work,
                   size_t work_bytes, void* stream);

/* Network input rows (bf16): out[i*kpad + q] = 1 - 2 bit_q(keys[i]) for
 * q < num_qubits, extra[i*n_extra + q - num_qubits] for the next n_extra
 * columns (per-state features, e.g. the scaled diagonal energy; extra may be
 * NULL when n_extra = 0), 0 up to kpad. */
/* Fused inference of the 256-wide network: y [n x 2] (FP32) for n keys and
 * their n x n_extra features, from the transposed bf16 weights W1^T [256 x
 * kpad], W2^T [256 x 256], W3^T [16 x 256] (rows 2-15 zero) and the FP32
 * biases. One kernel, activations in registers; the same roundings and order
 * as the training path's forward pass. No counterpart in the reference. */
int sqdk_mlp_infer(const uint64_t* keys, int64_t n, int32_t num_qubits, const float* extra,
                   int32_t n_extra, int32_t kpad, const uint16_t* w1t, const float* b1,
                   const uint16_t* w2t, const float* b2, const uint16_t* w3t, const float* b3,
                   float* y, void* stream);
int sqdk_mlp_encode(const uint64_t* keys, int64_t n, int32_t num_qubits, const float* extra,
                    int32_t n_extra, int32_t kpad, uint16_t* out, void* stream);

/* op 0: out = gelu(z); op 1: out = g * gelu'(z); op 2: out = z; GELU is the
 * exact erf form. Writes bf16 and/or FP32 outputs (either may be NULL). */
int sqdk_mlp_act(int32_t op, int64_t n, const float* z, const float* g, uint16_t* out_bf16,
                 float* out_f32, void* stream);

/* Weighted loss of outputs y[i*ldy + {0,1}] (log-amplitude, sign logit)
 * against target[3i..3i+2] = (log-amplitude, sign probability, weight):
 * sum w [(y0 - a)^2 + lambda BCE(sigmoid(y1), p)] / wsum. Writes dL/dy
 * (n x 2, FP32 and bf16) and nblocks partial sums of the loss. */
int sqdk_mlp_loss(int64_t n, const float* y, int64_t ldy, const float* target, double wsum,
                  double lambda, float* dy_f32, uint16_t* dy_bf16, double* partial,
                  int32_t nblocks, void* stream);

/* out[j] = sum over rows i of x[i*cols + j] (FP64 accumulation, row order). */
int sqdk_colsum(int64_t n, int64_t cols, const float* x, float* out, void* stream);

/* One Adam step on FP32 weights w with gradient g and moments m, v (step >= 1),
 * refreshing the bf16 copy w_bf16 (may be NULL). */
int sqdk_adam(int64_t n, float* w, const float* g, float* m, float* v, uint16_t* w_bf16,
              float lr, float beta1, float beta2, float eps, int32_t step, void* stream);

/* ---- state-vector simulation -------------------------------------------- */

/* One gate of a rotation circuit in the symplectic form of CircuitXZ
 * (rqutils/svsim.py:41-49): the state maps to cos psi + i sin Z^z X^x psi,
 * out[i] = cos psi[i] + i sin (-1)^{popcount(i & z)} psi[i ^ x]. */
typedef struct sqdk_gate {
  uint64_t x;
  uint64_t z;
  double cos;
  double sin;
} sqdk_gate;

/* Apply gates[0..n_gates) (host array) in order to the 2^num_qubits
 * amplitudes `state` (device, complex128 for SQDK_C128 or complex64 for
 * SQDK_C64) in place. Replaces do_svsim (rqutils/svsim.py:96-122); the
 * initial state is the caller's (rqutils/svsim.py:103-104 builds the one-hot
 * on the host side here). Consecutive gates whose X masks fit one LDS tile
 * share one pass over the state; runs_out (may be NULL) receives the number
 * of passes. num_qubits <= 40; every mask must lie below 2^num_qubits. */
int sqdk_svsim(int32_t num_qubits, int32_t precision, void* state, const sqdk_gate* gates,
               int64_t n_gates, int32_t* runs_out, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* SQDK_H */
