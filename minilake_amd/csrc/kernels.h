// Launch wrappers for the gfx950 kernels (kernels.hip) and the small structs
// they take by value.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mlgpu.h"

namespace mlg {

struct AccSlotList {
  int32_t kind[MAX_AGG_SLOTS];
};

// accumulator arrays moved by a rehash (old -> new slot)
struct AccMove {
  int32_t n;
  const void* src[MAX_AGG_SLOTS];
  void* dst[MAX_AGG_SLOTS];
};

// FIN_RANK: the state is the byte-order rank of a dictionary entry (MIN/MAX
// of strings); the output is the code of an entry with that rank.
enum FinishKind : int32_t { FIN_COPY = 0, FIN_AVG = 1, FIN_I32 = 2, FIN_RANK = 3 };

// hash aggregate output: key columns then one column per aggregate
struct EmitPlan {
  int32_t nkeys;
  int32_t key_nullable;
  int32_t key_f64;  // bit k: key k is FLOAT64 (NaN key words decode to the canonical NaN)
  int32_t key_width[MAX_KEYS];
  void* key_out[MAX_KEYS];
  uint8_t* key_valid[MAX_KEYS];
  int32_t nout;
  int32_t finish[MAX_AGG_SLOTS];
  int32_t out_width[MAX_AGG_SLOTS];
  const void* state[MAX_AGG_SLOTS];        // value state (sum / min / max / count)
  const void* count_state[MAX_AGG_SLOTS];  // valid-input count (AVG divisor, SUM/MIN/MAX seen)
  void* out[MAX_AGG_SLOTS];
  uint8_t* out_valid[MAX_AGG_SLOTS];
  const uint32_t* rank_code[MAX_AGG_SLOTS];  // FIN_RANK: entry code of each rank
  uint64_t rank_n[MAX_AGG_SLOTS];            // FIN_RANK: number of ranks
};

enum SortKeyKind : int32_t { SK_INT = 0, SK_F64 = 1, SK_UNSIGNED = 2 };
constexpr int MAX_SORT_KEYS = 4;
constexpr int TOPN_TILE = 1024;

struct SortKeys {
  int32_t nkeys;
  int32_t nwords;
  Prog prog[MAX_SORT_KEYS];
  int32_t kind[MAX_SORT_KEYS];
  int32_t asc[MAX_SORT_KEYS];
  int32_t nulls_first[MAX_SORT_KEYS];
  int32_t nullable[MAX_SORT_KEYS];
};

void launch_pipeline(const Plan& P, int grid, int block, size_t lds, hipStream_t s);
void launch_direct_reduce(const uint64_t* partials, int nblocks, int ncells, int nacc, const AccSlotList& accs,
                          uint64_t* out, hipStream_t s);
// wave-queued pipeline (flow_kernel) and the kernels around it
void launch_flow(const Plan& P, int grid, int block, size_t lds, hipStream_t s);
int flow_blocks_per_cu(int sink, bool fast, int block, size_t lds);
void launch_scan_counts(const uint32_t* counts, int64_t n, uint64_t* offsets, hipStream_t s);
void launch_materialize(const Plan& P, int sink, const uint64_t* offsets, int64_t nwaves, hipStream_t s);
void launch_fill_u64(uint64_t* p, uint64_t v, int64_t n, hipStream_t s);
// LSB-first validity bitmap (one bit per row) <-> one byte per row (1 = valid)
void launch_bits_to_bytes(const uint64_t* bits, int64_t n, uint8_t* out, hipStream_t s);
void launch_bytes_to_bits(const uint8_t* valid, int64_t n, uint64_t* bits, hipStream_t s);
// MLB1 batch bytes (byte-granular stores: section offsets are unaligned)
void launch_null_count(const uint8_t* valid, const uint32_t* sel, int64_t n, uint64_t* out, hipStream_t s);
void launch_mlb1_str_len(const uint32_t* codes, const uint32_t* sel, int64_t n, const uint32_t* dict_off,
                         uint32_t ndict, uint32_t* len, hipStream_t s);
void launch_mlb1_pack_fixed(const void* src, int width, bool as_bool, const uint32_t* sel, int64_t n, uint8_t* dst,
                            hipStream_t s);
void launch_mlb1_pack_validity(const uint8_t* valid, const uint32_t* sel, int64_t n, uint8_t* dst, hipStream_t s);
void launch_mlb1_pack_strings(const uint32_t* codes, const uint32_t* sel, int64_t n, const uint32_t* dict_off,
                              const uint8_t* dict_bytes, uint32_t ndict, const uint32_t* offs, uint8_t* dst,
                              hipStream_t s);
void launch_gather(const void* src, int width, const int64_t* idx, int64_t n, void* dst, const uint8_t* src_valid,
                   uint8_t* dst_valid, hipStream_t s);
void launch_join_insert(const DevJoinTable& T, const uint64_t* keys, const uint8_t* key_null, int64_t n,
                        uint64_t* bloom, hipStream_t s);
void launch_join_unique(const DevJoinTable& T, uint32_t* dup_flag, hipStream_t s);
// partitioned build (see join_part_build_kernel)
void launch_join_part_hist(const DevJoinTable& T, const uint64_t* keys, const uint8_t* key_null, int64_t n,
                           int nblocks, int64_t rows_per_block, int nparts, uint32_t* hist, hipStream_t s);
void launch_join_part_scatter(const DevJoinTable& T, const uint64_t* keys, const uint8_t* key_null, int64_t n,
                              int nblocks, int64_t rows_per_block, int nparts, const uint32_t* offsets,
                              uint32_t* srow, uint64_t* skeys, hipStream_t s);
void launch_join_part_build(const DevJoinTable& T, int nparts, const uint32_t* offsets, int nblocks,
                            const uint32_t* srow, const uint64_t* skeys, uint64_t* bloom, uint32_t* flags,
                            size_t lds, hipStream_t s);
void launch_strided_u32(const uint32_t* src, int64_t stride, int64_t n, uint32_t* dst, hipStream_t s);
// dense join tables: key range; presence bits (flag bit 0 = repeated key, bit 2 = key out of range);
// ranks (offsets[ne] = present keys); rows moved into key order
// Key word r of these kernels is keys[r * nkw] (the first of a row's words).
// Rows: n, or *n_dev when given (n then bounds it). mm[0..1] = min/max, mm[2] =
// the row count, mm[3] = the rows with a non-NULL key; partials holds 3 words
// per workgroup; `done` is a zeroed counter word (left zeroed); clear[0..1] are set to 0
void launch_key_minmax(const uint64_t* keys, int nkw, const uint8_t* key_null, int64_t n, const uint64_t* n_dev,
                       int64_t* partials, int nparts, int64_t* mm, uint32_t* done, uint32_t* clear, hipStream_t s);
// (flag bit 8: with nkw == 2, a second key word outside [0, 255])
void launch_dense_mark(const uint64_t* keys, int nkw, const uint8_t* key_null, int64_t n, int64_t kmin,
                       uint64_t range, uint64_t* bits, uint32_t* flag, hipStream_t s);
// Bloom bits of every non-NULL row's whole key, in the word of its first key
// (T: nkw; bloom holds ceil(range / 8) words, zeroed)
void launch_dense_bloom(const DevJoinTable& T, const uint64_t* keys, const uint8_t* key_null, int64_t n,
                        int64_t kmin, uint64_t range, uint64_t* bloom, hipStream_t s);
// rank[w] = set bits in words below w, rank[nw] = all set bits
void launch_dense_rank(const uint64_t* bits, int64_t nw, uint32_t* rank, uint32_t* block_sums, hipStream_t s);
// Offset-indexed dense tables (rows at their key offset, then the NULL-key
// rows from `range` on): row numbers in key order, given the rank of the
// presence bits (launch_dense_rank; rank[(range + 63) / 64] = keyed rows).
// translate: rows[i * stride] for i < n, in place; compact: every held
// position of a column (npos = range + NULL-key rows) moved to its row.
void launch_dense_translate(int64_t* rows, int64_t stride, int64_t n, const uint64_t* bits, const uint32_t* rank,
                            uint64_t range, hipStream_t s);
void launch_dense_compact(const void* src, int width, const uint8_t* src_valid, int64_t npos, const uint64_t* bits,
                          const uint32_t* rank, uint64_t range, void* dst, uint8_t* dst_valid, hipStream_t s);
struct DensePermute {
  int32_t nkw;                // key words per row; words 1.. go to dst_keys
  int32_t keys32;             // dst_keys holds u32 (the low halves) instead of u64
  uint64_t* dst_keys;
  uint8_t* dst_keys8;         // or null: the low byte of key word 1 (nkw == 2), the table's
                              // narrow copy when every second key word is in [0, 255]
  int32_t ncols;
  int32_t width[MAX_COLS];
  const void* src[MAX_COLS];
  void* dst[MAX_COLS];
  const uint8_t* src_valid[MAX_COLS];
  uint8_t* dst_valid[MAX_COLS];
  int64_t held;               // rows the destinations hold (0: the source's row count)
};
// Rows into key order; rank and present null for a table whose rows sit at
// their key's offset (a full table, where every key of the range is present,
// or an offset-indexed one), the NULL-key rows from `range` on.
// A first-key column rebuilt at row positions: dst[p] = kmin + p below
// `range`, 0 for the NULL-key rows after (width 4 or 8).
void launch_dense_key_fill(void* dst, int width, int64_t held, uint64_t range, int64_t kmin, hipStream_t s);
// Presence bits and rows of a table held at key offsets in one pass (M.held
// positions; flag bits as launch_dense_mark; *null_next from 0).
void launch_dense_place(const DensePermute& M, const uint64_t* keys, const uint8_t* key_null, int64_t n, int64_t kmin,
                        uint64_t range, uint64_t* bits, uint32_t* flag, uint32_t* null_next, hipStream_t s);
// mm = {min, max, rows, non-NULL rows} from nparts {min, max, count} partials
// of a join build's row kernel (rows = min(n, *n_dev)); clears clear[0..1].
void launch_key_range_reduce(const int64_t* partials, int nparts, int64_t n, const uint64_t* n_dev, int64_t* mm,
                             uint32_t* clear, hipStream_t s);
void launch_dense_permute(const DensePermute& M, const uint64_t* keys, const uint8_t* key_null, int64_t n,
                          int64_t kmin, uint64_t range, const uint64_t* bits, const uint32_t* rank,
                          const uint32_t* present, uint32_t* null_next, hipStream_t s);
// One-workgroup dense build of a small table (see small_dense_kernel); bits
// and rank hold small_dense_words() and small_dense_words() + 1 entries;
// status = {rows, min, range, flags (bit 0 repeated key, bit 1 not dense,
// SMALL_DENSE_FULL: dense and every key of the range present)}.
constexpr int64_t SMALL_DENSE_FULL = 256;
void launch_small_dense(const DensePermute& M, const uint64_t* keys, const uint8_t* key_null, int64_t cap,
                        const uint64_t* count_dev, uint64_t* bits, uint32_t* rank, int64_t* status, hipStream_t s);
int small_dense_words();
// ------------------------------------------------- partitioned aggregation
// Records of a hash aggregate, in partition order: key word w of record i at
// keys[w * n + i], argument a at args[a * n + i], argument validity bits at
// valid[i] (bit a) when any argument can be NULL.
struct AggRecords {
  int32_t nkeys;              // key columns (values materialized with validity)
  int32_t nkw;                // key words: nkeys (+ 1 null-mask word)
  int32_t nargs;
  int32_t key_nullable;
  const uint64_t* key_val[MAX_KEYS];
  const uint8_t* key_ok[MAX_KEYS];   // u8 validity or null
  const uint64_t* arg_val[MAX_AGG_SLOTS];
  const uint8_t* arg_ok[MAX_AGG_SLOTS];
};

// One LDS table per partition: the geometry and accumulator plan.
struct PartAggPlan {
  int32_t S;                  // slots per partition (power of two)
  int32_t nkw, nargs, nacc;
  uint64_t mask;              // table capacity - 1 (capacity = partitions * S)
  AccSlot acc[MAX_AGG_SLOTS];
  uint64_t* slots;            // final table: [slot][1 + nkw]
  void* acc_data[MAX_AGG_SLOTS];
  uint32_t* flags;            // [0..1] group count (u64), [2] overflow bits, [3] filled partitions,
                              // [4..5] projected groups of the filled partitions (u64)
};

// Groups an LDS partition table of S slots takes before it counts as full:
// 3/4 (linear probes for a new key stay short, and a table sized with the
// 2x margin sits near 1/2 full).
__host__ __device__ constexpr uint32_t part_fill_limit(uint32_t S) { return S - S / 4u; }

struct PartChunk {
  uint32_t part, lo, hi, partial;  // partial = index of its partial table, or ~0 (final)
};

void launch_agg_part_hist(const AggRecords& R, int64_t n, uint64_t mask, int part_shift, int nblocks,
                          int64_t rows_per_block, int nparts, uint32_t* hist, hipStream_t s);
void launch_agg_part_scatter(const AggRecords& R, int64_t n, uint64_t mask, int part_shift, int nblocks,
                             int64_t rows_per_block, int nparts, const uint32_t* offsets, uint64_t* rkeys,
                             uint64_t* rargs, uint32_t* rvalid, hipStream_t s);
void launch_agg_part_chunks(const PartAggPlan& A, const PartChunk* chunks, int nchunks, const uint64_t* rkeys,
                            const uint64_t* rargs, const uint32_t* rvalid, int64_t n, uint64_t* partial, hipStream_t s);
void launch_agg_part_merge(const PartAggPlan& A, const uint32_t* parts, const uint32_t* first, int nparts,
                           const uint64_t* partial, hipStream_t s);
size_t agg_part_lds(const PartAggPlan& A);
// exclusive scan of n u32 counts (any n < 2^32 total); out[n] = total
void launch_scan_u32(const uint32_t* in, int64_t n, uint32_t* out, uint32_t* block_sums, hipStream_t s);
int64_t scan_u32_blocks(int64_t n);
void launch_agg_rehash(const uint64_t* old_slots, uint64_t old_cap, int nkw, uint64_t* new_slots, uint64_t new_mask,
                       const AccMove& mv, hipStream_t s);
void launch_slot_flags(const uint64_t* slots, uint64_t cap, int nkw, uint8_t* flags, hipStream_t s);
void launch_agg_emit(const uint64_t* slots, int nkw, const int64_t* idx, int64_t n, const EmitPlan& E, hipStream_t s);
int64_t compact_blocks(int64_t n);
void launch_compact(const uint8_t* flags, int64_t n, uint32_t* block_counts, uint64_t* offsets, const uint32_t* sel,
                    void* out, int out_u32, hipStream_t s);
void launch_sort_keys(const Plan& P, const SortKeys& K, int64_t n, uint64_t* words, hipStream_t s);
void launch_topn_tiles(const uint64_t* in, int64_t n, int nw, int keep, uint64_t* out, hipStream_t s);
// The first `keep` (at most TOPN_RANK_LIMIT) entries of each TOPN_RANK_TILE-entry tile, by rank.
constexpr int TOPN_RANK_TILE = 256;
constexpr int TOPN_RANK_SPLIT = 4;  // threads counting one entry's rank
constexpr int TOPN_RANK_LIMIT = 32;
constexpr int MAX_SORT_WORDS = 6;  // key words + the row index
void launch_topn_rank(const uint64_t* in, int64_t n, int nw, int keep, uint64_t* out, hipStream_t s);
void launch_topn_rows(const uint64_t* words, int nw, int64_t n, int64_t* rows_out, hipStream_t s);
// Full sort of n entries of nw + 1 words (see merge_pass_kernel): a and b
// hold sort_padded_entries(n) entries each; returns the one with the result.
int64_t sort_padded_entries(int64_t n);
int64_t sort_split_entries(int64_t n);  // i64 entries of the merge splitter array
uint64_t* launch_sort_words(const uint64_t* words, int64_t n, int nw, uint64_t* a, uint64_t* b, int64_t* split,
                            hipStream_t s);

}  // namespace mlg
