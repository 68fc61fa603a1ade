// Shared host/device definitions of the gfx950 execution path.
//
// Plans reach the GPU as a `Plan` kernel argument (uniform, scalar-loaded)
// plus an instruction array in device memory. Expression programs are typed
// register code: each instruction names the stack slot it writes (its depth),
// so the interpreter never indexes registers dynamically.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mlg {

constexpr int MAX_DEPTH = 6;      // expression stack slots held in registers
constexpr int MAX_STAGES = 8;     // row-index stages: 0 = source, k = k-th probe
constexpr int MAX_STEPS = 12;     // filters + probes in one fused pipeline
constexpr int MAX_KEYS = 4;       // key columns (one 64-bit word each)
constexpr int MAX_AGG_SLOTS = 24; // accumulator state slots per aggregate sink
constexpr int MAX_COLS = 32;      // materialized columns (collect / join payload)
constexpr int MAX_LEVELS = 6;     // LDS queue levels of the wave-queued pipeline
constexpr int QCAP = 128;         // entries per queue level: one wave's pop leaves < 64,
                                  // one push adds <= 64, so 127 is the most ever held
#ifndef ML_FAST_U
#define ML_FAST_U 4
#endif
constexpr int FAST_U = ML_FAST_U;  // source rows per lane in a streaming first probe
// level 1 behind it holds < 64 + FAST_U * 64 entries
constexpr int QCAP_FAST = FAST_U <= 3 ? 256 : FAST_U <= 7 ? 512 : 1024;
static_assert(QCAP_FAST >= 64 * (FAST_U + 1), "level-1 queue too small");

// MiniLake's hash constants (crates/minilake-hashtable/src/hash.rs:12-17).
constexpr uint64_t HSEED = 0x243F6A8885A308D3ull;
constexpr uint64_t HMUL = 0x9E3779B97F4A7C15ull;
constexpr uint64_t HMUL2 = 0xBF58476D1CE4E5B9ull;
constexpr uint64_t HNULL = 0x7C159E3779B97F4Aull;

__host__ __device__ inline uint64_t fold_mul(uint64_t a, uint64_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return (a * b) ^ __umul64hi(a, b);
#else
  unsigned __int128 r = (unsigned __int128)a * b;
  return (uint64_t)r ^ (uint64_t)(r >> 64);
#endif
}
__host__ __device__ inline uint64_t hash_u64(uint64_t x) { return fold_mul(x ^ HSEED, HMUL); }
__host__ __device__ inline uint64_t hash_combine(uint64_t h, uint64_t v) { return fold_mul(h ^ v, HMUL2); }

// Hash of key words for this library's own device tables (join and aggregate
// tables, Bloom filters). Their layout is internal, so they use a cheaper mix
// than MiniLake's hash (one 64-bit multiply per key word instead of two
// 64x64->128-bit products, which cost ~16 VALU each on a 32-bit ALU);
// hash_columns still returns MiniLake's hash.
constexpr uint64_t TSEED = 0x632BE59BD9B4E019ull;
__host__ __device__ inline uint64_t tmix(uint64_t x) {
  x *= 0x9E3779B97F4A7C15ull;
  return x ^ (x >> 32);
}
__host__ __device__ inline uint64_t table_hash(const uint64_t* w, int n) {
  uint64_t h = tmix(w[0] ^ TSEED);
  for (int k = 1; k < n; k++) h = tmix(h ^ w[k]);
  return h;
}
// Bloom hash of a table hash: its top bits pick the 3 bits, bits 18.. the word.
__host__ __device__ inline uint64_t bloom_mix(uint64_t h) { return h * 0xD6E8FEB86659FD93ull; }

// Group key words of NaN keys: sign + exponent + quiet bit set, the low 51
// bits a number unique to the row. Results decode any NaN key word to the
// canonical quiet NaN.
constexpr uint64_t NAN_KEY_TAG = 0xFFF8000000000000ull;
constexpr uint64_t NAN_KEY_IDS = 0x0007FFFFFFFFFFFFull;
constexpr uint64_t CANONICAL_NAN = 0x7FF8000000000000ull;
__host__ __device__ inline bool f64_bits_nan(uint64_t w) {
  return ((w >> 52) & 0x7FFull) == 0x7FFull && (w & 0x000FFFFFFFFFFFFFull) != 0;
}

// ------------------------------------------------------------------ opcodes
enum Opcode : uint16_t {
  // push (write slot `depth`)
  OP_LOAD_I32 = 0,  // sign-extended
  OP_LOAD_U32,      // dictionary codes
  OP_LOAD_I64,
  OP_LOAD_U8,       // booleans (normalized to 0/1)
  OP_CONST,
  // fused push forms (a peephole of LOAD CONST CMP and LOAD LOOKUP):
  OP_CMPC,          // column imm (at `stage`, load kind flags>>13 & 3 = I32/U32/I64/U8) compared
                    // with the constant imm2: flags>>8 & 7 = EQ,NE,LT,LE,GT,GE, bit 11 = f64
  OP_LOOKUPC,       // table imm2 indexed by the u32 column imm (a dictionary predicate)
  // unary on slot depth-1
  OP_NOT,
  OP_NEG_I32,
  OP_NEG_I64,
  OP_NEG_F64,
  OP_I2F,           // i64 (or sign-extended i32 / bool) -> f64
  OP_F2I64,         // saturating, NaN -> 0 (Rust `as`)
  OP_F2I32,
  OP_I64_TO_I32,    // truncate, sign-extend
  OP_LOOKUP_U8,     // slot = table[slot] (dictionary predicate per entry)
  OP_LOOKUP_U64,    // slot = table[slot] (dictionary hash / rank per entry)
  OP_KEY_F64,       // group key word: -0.0 -> +0.0; a NaN becomes a word no other row has
                    // (a NaN tag plus a number drawn from the u64 counter at imm), so every
                    // NaN row is its own group, as f64 `==` never matches NaN (kernels/hash.rs:94)
  OP_HASH_F64,      // hash input word: 0 if value == 0.0 else bits (kernels/hash.rs:53)
  OP_JKEY_F64,      // join key word: -0.0 -> +0.0; a NaN key is NULL, so it never matches
  // binary: slots depth-2, depth-1 -> depth-2
  OP_ADD_I32, OP_SUB_I32, OP_MUL_I32, OP_MOD_I32,
  OP_ADD_I64, OP_SUB_I64, OP_MUL_I64, OP_MOD_I64,
  OP_ADD_F64, OP_SUB_F64, OP_MUL_F64, OP_DIV_F64, OP_MOD_F64,
  OP_EQ_I, OP_NE_I, OP_LT_I, OP_LE_I, OP_GT_I, OP_GE_I,
  OP_EQ_F, OP_NE_F, OP_LT_F, OP_LE_F, OP_GT_F, OP_GE_F,
  OP_AND, OP_OR,
  OP_COUNT
};

// The first 8 bytes are fetched as one little-endian word on the device:
// keep this field order.
struct Instr {
  uint16_t op;
  uint8_t depth;   // stack depth before the instruction
  uint8_t stage;   // LOAD: which row index
  uint32_t flags;  // CONST: 1 = valid
  uint64_t imm;    // LOAD: data pointer; CONST: value bits; LOOKUP: table pointer
  uint64_t imm2;   // LOAD: u8 validity pointer or 0
};

// Instructions that read a column at their stage.
__host__ __device__ inline bool reads_column(uint32_t op) {
  return op < OP_CONST || op == OP_CMPC || op == OP_LOOKUPC;
}
// The plain load opcode a column-reading instruction uses.
__host__ __device__ inline uint32_t column_load_op(uint32_t op, uint32_t flags) {
  if (op == OP_CMPC) return (flags >> 13) & 3u;
  if (op == OP_LOOKUPC) return OP_LOAD_U32;
  return op;
}

struct Prog {
  int32_t first = 0;     // index into the plan's instruction array
  int16_t count = 0;
  int16_t nullable = 0;  // 1 if the program can produce NULL (tracks validity bits)
};

// ---------------------------------------------------------------- hash tables
// Join table: open addressing, linear probing; slot = header word then key
// words. header = (tag << 32) | (build_row + 1), 0 = empty. The slots are cut
// into partitions of 2^part_shift slots (built one per workgroup in LDS); a
// probe starts at slot hash & mask and wraps inside its partition. A table
// that could not be partitioned has one partition (part_shift = log2 capacity).
// The blocked Bloom filter has 2^bloom_shift words per partition.
struct DevJoinTable {
  uint64_t* slots = nullptr;
  uint64_t mask = 0;       // capacity - 1
  uint64_t wrap = 0;       // slots per partition - 1
  int32_t nkw = 1;         // key words
  int32_t unique = 1;      // every key distinct: stop at the first match
  int32_t part_shift = 0;  // log2 slots per partition
  int32_t bloom_shift = 0; // log2 Bloom words per partition
  const uint64_t* bloom = nullptr;  // blocked Bloom filter (one u64 per block) or null
  uint64_t bloom_mask = 0; // Bloom words per partition - 1
  // Dense layout (the first key column is an integer, unique, its values
  // over a small range): the table's rows are in order of the first key word
  // k, so the row of k is the number of present first keys below it. With
  // i = k - dense_min < dense_range, k is present iff bit i % 64 of
  // dense_bits[i / 64] is set, and its row is dense_rank[i / 64] +
  // popcount(dense_bits[i / 64] & ((1 << i % 64) - 1)). Further key words
  // must equal entry row * (nkw - 1) + j - 1 of dense_keys: u64 words, or u32
  // (dense_keys32) when every further key column is 4 bytes wide, widened by
  // sign (bit j of dense_keys_sx) or zero. A Bloom filter can sit in front:
  // word i >> DENSE_BLOOM_SHIFT (a word per 8 consecutive first-key values,
  // so rows in key order set runs of one word), bits from the whole key's
  // hash. Slots are unused. dense_full: every key of the range is present,
  // so the row of k is i itself and probes read neither bits nor ranks (nor
  // a Bloom filter: the further key words are one load away).
  int32_t dense = 0;
  int32_t dense_full = 0;
  int64_t dense_min = 0;
  uint64_t dense_range = 0;
  const uint64_t* dense_bits = nullptr;  // one bit per key of the range
  const uint32_t* dense_rank = nullptr;  // present keys below each word
  const uint64_t* dense_keys = nullptr;  // key words 1.. of each row (nkw > 1)
  int32_t dense_keys32 = 0;
  int32_t dense_keys8 = 0;  // nkw == 2 and every second key word in [0, 255]: one byte per row
  uint32_t dense_keys_sx = 0;
};

constexpr int DENSE_BLOOM_SHIFT = 3;

// Key word j (>= 1) of dense-table row `row`.
__host__ __device__ inline uint64_t dense_key_word(const DevJoinTable& T, uint64_t row, int j) {
  if (T.dense_keys8) return ((const uint8_t*)T.dense_keys)[row];
  const uint64_t at = row * (uint64_t)(T.nkw - 1) + (uint64_t)(j - 1);
  if (!T.dense_keys32) return T.dense_keys[at];
  const uint32_t v = ((const uint32_t*)T.dense_keys)[at];
  return ((T.dense_keys_sx >> j) & 1u) ? (uint64_t)(int64_t)(int32_t)v : (uint64_t)v;
}

// next slot of a linear probe (inside the partition)
__host__ __device__ inline uint64_t join_step(const DevJoinTable& T, uint64_t pos) {
  return (pos & ~T.wrap) | ((pos + 1) & T.wrap);
}
// Bloom word of a key with table hash h (g = bloom_mix(h) picks the word and bits)
__host__ __device__ inline uint64_t bloom_word(const DevJoinTable& T, uint64_t h, uint64_t g) {
  return (((h & T.mask) >> T.part_shift) << T.bloom_shift) | ((g >> 18) & T.bloom_mask);
}
__host__ __device__ inline uint64_t bloom_bits(uint64_t g) {
  return (1ull << ((g >> 46) & 63)) | (1ull << ((g >> 52) & 63)) | (1ull << (g >> 58));
}

// Aggregate table: header word per slot (0 empty, 1 being written,
// otherwise (hash & ~3) | 2), key words AoS, accumulators SoA by slot.
enum AccKind : int32_t {
  ACC_SUM_F64 = 0,   // f64 += v
  ACC_SUM_I64 = 1,   // wrapping i64 += v
  ACC_COUNT = 2,     // i64 += 1 (when the argument is valid)
  ACC_MIN_I64 = 3,
  ACC_MAX_I64 = 4,
  ACC_MIN_F64 = 5,
  ACC_MAX_F64 = 6,
};
// Identity element of each accumulator state.
__host__ __device__ inline uint64_t acc_init_bits(int kind) {
  switch (kind) {
    case ACC_MIN_I64: return 0x7FFFFFFFFFFFFFFFull;
    case ACC_MAX_I64: return 0x8000000000000000ull;
    case ACC_MIN_F64: return 0x7FF0000000000000ull;  // +inf
    case ACC_MAX_F64: return 0xFFF0000000000000ull;  // -inf
    default: return 0;
  }
}

struct AccSlot {
  int32_t kind;
  int32_t arg;       // index of the argument program (-1 = rows counter)
};

enum SinkKind : int32_t {
  SINK_NONE = 0,
  SINK_HASH_AGG = 1,
  SINK_DIRECT_AGG = 2,
  SINK_JOIN_BUILD = 3,
  SINK_COLLECT = 4,
  SINK_COUNT_PAIRS = 5,
  SINK_PROJECT = 6,   // write every column at its physical row
  SINK_HASH = 7,      // hash_columns: one hash per active row
  SINK_FLAGS = 8,     // filter: one 0/1 flag per active row
  SINK_ROWS = 9,      // wave-queued pipeline: surviving rows' stage indices, per-wave regions
};

enum StepKind : int32_t { STEP_FILTER = 0, STEP_PROBE = 1 };

struct Step {
  int32_t kind;
  int32_t stage;       // PROBE: stage whose row index it sets
  int32_t table;       // PROBE: index into Plan::tables
  int32_t nkeys;
  Prog pred;           // FILTER
  Prog keys[MAX_KEYS]; // PROBE: key word programs
};

constexpr int MAX_ATOMS = 4;
struct FilterAtom {
  const void* col;   // stage-0 column
  uint64_t imm2;     // CMPC: the constant; LOOKUPC: the table
  uint32_t op;       // OP_CMPC or OP_LOOKUPC
  uint32_t flags;    // the instruction's flags
};

struct Plan {
  // source rows: i in [begin, end) of the selection (or of the rows)
  const uint32_t* sel;
  int64_t begin;
  int64_t end;
  const Instr* ins;
  int32_t nins;              // instructions in `ins`

  int32_t nsteps;
  int32_t expand_step;  // index of the one probe that may match many rows, or -1
  Step steps[MAX_STEPS];
  DevJoinTable tables[MAX_STAGES];

  // ---- sink
  int32_t sink;
  int32_t nkeys;             // AGG / JOIN_BUILD: key programs
  Prog keys[MAX_KEYS];
  int32_t key_nullable;      // AGG: keys may be NULL (adds a null-mask word)
  int32_t key_prehashed;     // HASH: bit k set = key k already is a 64-bit hash
  uint64_t* hash_out;        // HASH: one hash per active row
  uint8_t* flag_out;         // FILTER: one predicate result per active row

  // hash aggregate
  uint64_t* agg_slots;       // header + key words, stride 1 + agg_nkw
  int32_t agg_nkw;
  int32_t nacc;
  uint64_t agg_mask;
  uint64_t agg_wrap;         // linear probing wraps within agg_wrap + 1 slots (a partition or the table)
  uint64_t* agg_count;       // number of groups (device counter)
  uint32_t* agg_overflow;    // bit 0: the table passed its fill limit; bit 1: a row found no slot
  uint64_t agg_fill_limit;
  // Pre-sized table (hint / learned sizing): passing the fill limit means the
  // estimate was too small, which counts as a lost row (bit 1), and once that
  // is set every later row stops at once (the caller reruns the batch or
  // fails), so a far too small estimate never fills the table up to long
  // probe chains. 0: a grow-and-rehash table (reserved before each batch).
  uint64_t agg_presized;
  AccSlot acc[MAX_AGG_SLOTS];
  int32_t nargs;
  Prog args[MAX_AGG_SLOTS];  // argument programs
  uint32_t arg_slots[MAX_AGG_SLOTS];  // bit s of arg_slots[a]: slot s accumulates argument a
  uint32_t count_slots;      // bit s: slot s counts rows (no argument)
  uint32_t arg_chain;        // bit a: args[a] begins with all of args[a-1]'s code, and a
                             // 4-row evaluation continues from that result
  int32_t src_batch;         // 1: every key and argument program needs at most 4 stack
                             // slots, so a level-0 sink may evaluate several rows at once
  // stage-0 columns a batched level-0 sink reads, loaded for the round's rows
  // before the round's filter runs; an instruction reading column j carries
  // j + 1 in flags bits 24..27 (ignored by the one-row interpreter)
  int32_t sp_n;
  int32_t sp_op[8];          // OP_LOAD_* kind of column j
  const void* sp_col[8];
  void* acc_data[MAX_AGG_SLOTS];  // SoA arrays (f64 or i64) of capacity slots

  // direct (small-domain) aggregate: key k contributes (word * dim_stride[k])
  int64_t dim_stride[MAX_KEYS];
  int64_t dim_size[MAX_KEYS];
  int32_t ngroups;
  int32_t direct_copies;     // flow kernel: copies of the cells in LDS (a power of two)
  int32_t direct_stride;     // words between copies (odd, >= ngroups * nacc)
  void* direct_partials;     // [grid][ngroups][nacc] 8-byte words

  // join build / collect: materialized columns
  int32_t ncols;
  Prog cols[MAX_COLS];
  int32_t col_width[MAX_COLS];
  void* col_out[MAX_COLS];
  uint8_t* col_valid[MAX_COLS];  // u8 validity or null
  uint64_t* key_out;         // JOIN_BUILD: key words [rows][nkeys]
  uint8_t* key_null;         // JOIN_BUILD: 1 if any key is NULL
  int64_t* key_partials;     // JOIN_BUILD: per workgroup {min, max, count} of key word 0 over non-NULL rows, or null
  uint64_t* out_count;       // rows written so far (device counter)
  uint64_t out_capacity;
  uint32_t* out_overflow;
  int32_t out_identity;      // JOIN_BUILD: every source row is output, at its own index (no counter)
  int64_t* pair_out;         // COLLECT with row ids: [rows][nstages]
  int32_t pair_stages;

  // ---- wave-queued execution (flow_kernel). The steps are cut into
  // segments; between two segments every wave keeps the stage row indices of
  // its surviving rows in an LDS queue, and runs the next segment only on
  // full waves of 64 queued rows, so selective filters and probes leave no
  // lanes idle further down the chain. Each wave owns `chunk` source rows.
  int32_t nseg;                          // segments (queue levels 1 .. nseg-1)
  int32_t qwords;                        // u32 words of one wave's queues
  // (32-bit fields: they are read with scalar loads, which have no sub-dword form)
  int32_t seg_end[MAX_LEVELS + 1];       // segment s = steps [seg_end[s-1], seg_end[s])
  int32_t qoff[MAX_LEVELS + 1];          // level l queue: u32 offset in the wave's region
  int32_t qcap[MAX_LEVELS + 1];          // level l queue: entries (power of two)
  uint32_t live[MAX_LEVELS + 1];         // level l carries these stages' row indices (bit s = stage s)
  uint32_t tup_mask;                     // ROWS: stages recorded per row
  int64_t chunk;                         // source rows per wave
  uint32_t* tuples;                      // ROWS: [k][staging] row index of the k-th recorded stage
  uint64_t staging;                      // waves * chunk
  uint32_t* wave_counts;                 // ROWS / COUNT_PAIRS: rows produced per wave (flow), or per
                                         // thread (row-per-lane COUNT_PAIRS counting pass)
  uint64_t* thread_offsets;              // row-per-lane COLLECT: each thread's first output row
  // streaming first probe: step 0 probes a unique table with plain stage-0
  // column keys; level 0 then handles FAST_U rows per lane with independent
  // key, Bloom and slot loads in flight
  int32_t fast;
  int32_t fast_nkeys;
  uint16_t fast_op[2];
  const void* fast_col[2];
  // otherwise: stage-0 columns segment 0 reads, loaded ahead for FAST_U rows
  int32_t pf_n;
  uint16_t pf_op[2];
  const void* pf_col[2];
  // Hash aggregate: per-workgroup LDS cache of group accumulators, keyed by
  // global slot (ac_n entries, a power of two; 0 = off), at u64 offset ac_off
  // of the dynamic LDS. Rows of cached groups update LDS; the cache is merged
  // into the table once per workgroup, so hot groups cost no global atomics.
  int32_t ac_n;
  int32_t ac_off;
  // Segment 0 is one filter made only of fused column atoms (OP_CMPC /
  // OP_LOOKUPC on stage 0) joined by AND / OR / NOT: the atoms are evaluated
  // directly and the predicate is bit (atom outcomes) of atom_truth.
  int32_t natoms;            // 0 = not this form
  uint32_t atom_truth;       // bit m = predicate when atom i's outcome is bit i of m
  FilterAtom atoms[MAX_ATOMS];
};

// Sortable key words for TopN.
struct SortPlan {
  int32_t nwords;
};

}  // namespace mlg
