/*
 * minilake_gpu.h — C ABI of the MI355X (gfx950) execution path for MiniLake's
 * data-parallel core: filter/projection expression kernels, hash aggregate,
 * hash join build/probe, TopN, fused pipelines, batched fragments and learned
 * hash-table sizing.
 *
 * Everything crossing this boundary is plain C: POD structs, pointers, sizes
 * and int status codes. No exception or unwinding crosses it. Column data
 * pointers given to the GPU library are DEVICE pointers; dictionaries, plan
 * structs and expression programs live in HOST memory.
 *
 * The reference interfaces each group of entry points replaces are cited per
 * declaration (paths relative to the MiniLake repository root).
 *
 * Ownership rules (mirror crates/minilake-exec/src/pipeline.rs:32-79):
 *   - the caller owns inputs for the duration of a call;
 *   - results are allocated by the library and released with ml_result_free;
 *   - one ml_ctx (one HIP stream) per host thread; handles are not
 *     thread-safe, like LocalSink;
 *   - finalize is called once, after the last sink.
 */
#ifndef MINILAKE_GPU_H
#define MINILAKE_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ */
/* Status codes: one per MiniLakeError variant
 * (crates/minilake-core/src/error.rs:8-48), 0 = Ok.                  */
/* ------------------------------------------------------------------ */
typedef int32_t ml_status;
enum {
  ML_OK = 0,
  ML_ERR_PARSE = 1,
  ML_ERR_PLAN = 2,
  ML_ERR_EXECUTION = 3,
  ML_ERR_UNSUPPORTED = 4,
  ML_ERR_IO = 5,
  ML_ERR_PARQUET = 6,
  ML_ERR_RESOURCES_EXHAUSTED = 7,
  ML_ERR_CANCELLED = 8,
  ML_ERR_INTERNAL = 9,
  /* Not an error of the reference: a pre-sized hash table was too small.
   * Returned by ML_SIZING_HINT; the learned-sizing path grows or falls back
   * instead (see ML_SIZING_LEARNED). */
  ML_ERR_CAPACITY = 10
};

/* ------------------------------------------------------------------ */
/* Data types (crates/minilake-core/src/types.rs:15-28)                */
/* ------------------------------------------------------------------ */
enum {
  ML_TYPE_BOOLEAN = 0, /* 1 byte per value (Vec<bool>) */
  ML_TYPE_INT32 = 1,
  ML_TYPE_INT64 = 2,
  ML_TYPE_FLOAT64 = 3,
  ML_TYPE_DATE = 4, /* i32 days since 1970-01-01 */
  ML_TYPE_UTF8 = 5, /* dictionary-encoded: u32 code per row + ml_dict; every code,
                       NULL rows' included, is below the dictionary's len */
  ML_TYPE_NULL = -1 /* only for a NULL literal */
};

/* Dictionary of a UTF8 column (DictVec, crates/minilake-core/src/column.rs:126-131).
 * Host memory. Entry i is bytes[offsets[i] .. offsets[i+1]). */
typedef struct ml_dict {
  int32_t len;
  int32_t _pad;
  const uint32_t* offsets; /* len + 1 entries */
  const uint8_t* bytes;
} ml_dict;

/* One column (Column/ColumnData, crates/minilake-core/src/column.rs:135-200). */
typedef struct ml_column {
  int32_t type;
  int32_t _pad;
  const void* data;          /* device pointer, `len` values of the type's width */
  int64_t len;
  const uint64_t* validity;  /* device LSB-first bitmap (Bitmap, bitmap.rs:9-12),
                                ceil(len / 64) words; NULL = all valid */
  const ml_dict* dict;       /* UTF8 only */
} ml_column;

/* A batch (crates/minilake-core/src/batch.rs:25-29): equal-length columns plus
 * an optional ascending selection vector (batch.rs:14). */
typedef struct ml_batch {
  const ml_column* columns;
  int32_t num_columns;
  int32_t _pad;
  int64_t num_rows;
  const uint32_t* selection; /* device pointer, ascending; NULL = all rows */
  int64_t selection_len;
} ml_batch;

/* ------------------------------------------------------------------ */
/* Expressions: PhysicalExpr (crates/minilake-exec/src/expr/mod.rs:109-173)
 * flattened to postfix. Column indices refer to the row layout at the
 * point the expression is evaluated (see ml_operator).               */
/* ------------------------------------------------------------------ */
enum {
  ML_EXPR_COLUMN = 0,
  ML_EXPR_LITERAL = 1,
  ML_EXPR_BINARY = 2,
  ML_EXPR_NOT = 3,
  ML_EXPR_NEGATE = 4,
  ML_EXPR_CAST = 5
};
/* BinaryOp, same order as crates/minilake-exec/src/expr/mod.rs:20-47 */
enum {
  ML_OP_ADD = 0, ML_OP_SUB, ML_OP_MUL, ML_OP_DIV, ML_OP_MOD,
  ML_OP_EQ, ML_OP_NOTEQ, ML_OP_LT, ML_OP_LTEQ, ML_OP_GT, ML_OP_GTEQ,
  ML_OP_AND, ML_OP_OR
};

typedef struct ml_expr_node {
  int32_t kind;   /* ML_EXPR_* */
  int32_t op;     /* ML_OP_* for BINARY */
  int32_t type;   /* literal type / cast target */
  int32_t column; /* COLUMN index */
  int64_t i64;    /* INT32/INT64/DATE/BOOLEAN literal */
  double f64;     /* FLOAT64 literal */
  const char* str; /* UTF8 literal (host memory) */
  int32_t str_len;
  int32_t _pad;
} ml_expr_node;

typedef struct ml_expr {
  const ml_expr_node* nodes;
  int32_t num_nodes;
  int32_t _pad;
} ml_expr;

/* ------------------------------------------------------------------ */
/* Execution context (ExecConfig, crates/minilake-exec/src/context.rs:14-34) */
/* ------------------------------------------------------------------ */
typedef struct ml_exec_config {
  int64_t batch_rows;       /* device batch (morsel) size; default 2^21 */
  int32_t device;           /* HIP device ordinal */
  int32_t block_size;       /* threads per workgroup for streaming kernels: 0 = tuned, else 64, 128, 192 or 256 */
  int32_t grid_per_cu;      /* workgroups per CU for grid-stride kernels (0 = tuned) */
  int32_t use_bloom;        /* 1: join tables carry a blocked Bloom filter (-1 = auto) */
  void* stream;             /* hipStream_t to launch on; NULL = library-owned stream */
} ml_exec_config;

typedef struct ml_ctx ml_ctx;
typedef struct ml_result ml_result;
typedef struct ml_hashagg ml_hashagg;
typedef struct ml_join_build ml_join_build;
typedef struct ml_join_table ml_join_table;
typedef struct ml_topn ml_topn;

void ml_default_config(ml_exec_config* cfg);
ml_status ml_ctx_create(const ml_exec_config* cfg, ml_ctx** out);
void ml_ctx_destroy(ml_ctx* ctx);
/* Message of the last error raised on this context ("" if none). */
const char* ml_last_error(const ml_ctx* ctx);
/* Wait for all work queued on the context's stream. */
ml_status ml_ctx_synchronize(ml_ctx* ctx);
/* The context's hipStream_t. */
void* ml_ctx_stream(ml_ctx* ctx);

/* Per-launch statistics of fused pipeline kernels (the GPU analog of
 * OperatorMetrics rows_in / time, crates/minilake-exec/src/metrics.rs:13-40):
 * kernel time from HIP events recorded on the launch stream. */
typedef struct ml_launch_stat {
  int32_t sink;   /* 1 hash agg, 2 direct agg, 3 join build, 4 collect, ... */
  int32_t _pad;
  int64_t rows;   /* source rows the launch covered */
  double ms;      /* kernel duration */
} ml_launch_stat;
/* on: 0 off, 1 every launch, N > 1 only launches over at least N source rows
 * (so recording does not slow the small launches around a large one). */
ml_status ml_ctx_enable_stats(ml_ctx* ctx, int32_t on);
/* Synchronize, copy up to `cap` recorded launches to `out`, clear the record;
 * returns the number written. */
int32_t ml_ctx_take_stats(ml_ctx* ctx, ml_launch_stat* out, int32_t cap);

/* MemoryPool (crates/minilake-exec/src/memory.rs:22-157): a budget in bytes
 * for the context's pipeline-breaker state (hash-aggregate group tables, join
 * tables from ml_join_build_finalize until ml_join_table_free); 0 = unlimited.
 * A join table over budget is ML_ERR_RESOURCES_EXHAUSTED. Over budget, a hash aggregate returns
 * ML_ERR_RESOURCES_EXHAUSTED, or with spill != 0 writes its groups' partial
 * states to host memory in 16 hash partitions (aggregate/hash.rs:189-282) and
 * merges one partition at a time at finalize. Restarts the peak below from the
 * bytes reserved now. */
ml_status ml_ctx_set_memory_limit(ml_ctx* ctx, int64_t bytes, int32_t spill);
/* Bytes reserved now, the most reserved at once since the limit was last set
 * (or since the context was created), and the limit. */
void ml_ctx_memory_usage(const ml_ctx* ctx, int64_t* used, int64_t* peak, int64_t* limit);

/* Results: library-owned device columns (host copies on request). */
int64_t ml_result_num_rows(const ml_result* r);
int32_t ml_result_num_columns(const ml_result* r);
/* Column i of the result; its data and validity bitmap are device memory,
 * so a result can be the source batch of another pipeline. They are written
 * on the context's stream: read them on that stream (ml_ctx_stream) or after
 * ml_ctx_synchronize. */
const ml_column* ml_result_column(const ml_result* r, int32_t i);
/* Copy column i to `dst`, host or device memory (num_rows * width bytes). */
ml_status ml_result_copy_column(ml_ctx* ctx, const ml_result* r, int32_t i, void* dst);
/* Copy column i's validity as one byte per row (1 = valid) to host memory. */
ml_status ml_result_copy_validity(ml_ctx* ctx, const ml_result* r, int32_t i, uint8_t* dst);
void ml_result_free(ml_result* r);

/* Allocation-free variant of ml_filter. It has no counterpart in the
 * reference: it is for callers that cannot allocate device memory (work under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, results and errors as ml_filter, but it allocates nothing:
 * every device buffer of the call (flags, compaction counts and offsets,
 * expanded validity, program, lookup tables) comes from one static device
 * array of ML_FILTER_NOALLOC_SCRATCH_BYTES per GPU, held until the process
 * ends. Limits: at most ML_FILTER_NOALLOC_MAX_ROWS active rows, and the call's
 * buffers (1 byte per active row, 1 byte per row for each column with a
 * validity bitmap, 12 bytes per 1024 active rows, the program) must fit the
 * array; otherwise ML_ERR_RESOURCES_EXHAUSTED before any launch. Every call on
 * one GPU shares that GPU's array: calls that overlap on one GPU (several
 * streams or host threads at once) race; calls on one stream are ordered. */
#define ML_FILTER_NOALLOC_SCRATCH_BYTES (32 * 1024 * 1024)
#define ML_FILTER_NOALLOC_MAX_ROWS (1 << 24)
ml_status ml_filter_noalloc(ml_ctx* ctx, const ml_expr* predicate, const ml_batch* in,
                            uint32_t* out_sel, int64_t* out_len);

/* ------------------------------------------------------------------ */
/* Device Parquet scan (crates/minilake-storage/src/table.rs:59-240,
 * parquet_reader.rs:210-358, stats.rs:53-70). Footer and page headers are
 * parsed on the host (SNAPPY page bodies decompressed there); page bodies
 * are decoded on the device. Flat schemas, UNCOMPRESSED and SNAPPY chunks,
 * data pages v1, PLAIN and dictionary encodings, RLE definition levels;
 * other files return ML_ERR_UNSUPPORTED.                                 */
/* ------------------------------------------------------------------ */
typedef struct ml_parquet ml_parquet;

/* RowGroupMeta.stats (convert_stats, parquet_reader.rs:290-358): exact
 * bounds only; none for booleans and byte-array decimals; string bounds only
 * when valid UTF-8 (pointers valid while the handle lives). */
typedef struct ml_column_stats {
  int32_t type;           /* ml_type of the column */
  int32_t has_min, has_max, has_null_count;
  int64_t null_count;
  int64_t min_i64, max_i64;  /* INT32 / DATE / INT64 */
  double min_f64, max_f64;   /* FLOAT64 */
  const char* min_str;       /* UTF8 */
  int64_t min_len;
  const char* max_str;
  int64_t max_len;
} ml_column_stats;

enum { /* PruneOp (stats.rs:24-36) */
  ML_PRUNE_EQ = 0, ML_PRUNE_LT = 1, ML_PRUNE_LTEQ = 2, ML_PRUNE_GT = 3, ML_PRUNE_GTEQ = 4
};
/* PrunePredicate (stats.rs:38-47): column op literal (i64 for INT32 / INT64 /
 * DATE / BOOLEAN, f64 for FLOAT64, str for UTF8). */
typedef struct ml_prune_predicate {
  int32_t column;
  int32_t op;
  int32_t type;
  int32_t _pad;
  int64_t i64;
  double f64;
  const char* str;
  int64_t str_len;
} ml_prune_predicate;

/* Table::open (table.rs:59-134) over one file held in host memory; the bytes
 * must stay valid while the handle lives. */
ml_status ml_parquet_open(ml_ctx* ctx, const void* data, int64_t len, ml_parquet** out);
int32_t ml_parquet_num_columns(const ml_parquet* p);
int32_t ml_parquet_column_type(const ml_parquet* p, int32_t column);
const char* ml_parquet_column_name(const ml_parquet* p, int32_t column);
int32_t ml_parquet_column_nullable(const ml_parquet* p, int32_t column);
int32_t ml_parquet_num_row_groups(const ml_parquet* p);
int64_t ml_parquet_row_group_rows(const ml_parquet* p, int32_t row_group);
ml_status ml_parquet_column_stats(const ml_parquet* p, int32_t row_group, int32_t column, ml_column_stats* out);
/* Table::prune (table.rs:165-172): the row groups that may hold matching rows
 * (out_row_groups holds at least ml_parquet_num_row_groups entries). */
ml_status ml_parquet_prune(const ml_parquet* p, const ml_prune_predicate* preds, int32_t num_preds,
                           int32_t* out_row_groups, int32_t* out_n);
/* Table::read_row_group (table.rs:204-240) + read_column_chunk
 * (parquet_reader.rs:210-287): the projected columns of one row group as a
 * device result (strings dictionary-encoded; DATE, DECIMAL -> FLOAT64 as
 * value / 10^scale, FLOAT -> FLOAT64). */
ml_status ml_parquet_read_row_group(ml_parquet* p, int32_t row_group, const int32_t* projection,
                                    int32_t num_columns, ml_result** out);
void ml_parquet_free(ml_parquet* p);

/* ------------------------------------------------------------------ */
/* MLB1 batch bytes (crates/minilake-core/src/ipc.rs:1-179): the format of
 * aggregate spill files and of partial aggregates shipped between workers. */
/* ------------------------------------------------------------------ */
/* write_batch (ipc.rs:44-87) on the device: the batch's selection is applied,
 * dictionary columns are written as plain strings, and a column carries
 * validity words only when a kept row is NULL. *size receives the byte count.
 * With dst != NULL (device memory, cap bytes) the bytes are written there
 * (cap < size: ML_ERR_CAPACITY); the call returns when they are written. */
ml_status ml_mlb1_write(ml_ctx* ctx, const ml_batch* batch, void* dst, int64_t cap, int64_t* size);
/* read_batch (ipc.rs:111-170): one batch from `len` bytes of host memory into
 * a device result (strings dictionary-encoded); *consumed = bytes used. Fewer
 * than 4 bytes is a clean end of stream (*out = NULL, ML_OK); a batch cut
 * short is ML_ERR_IO; a bad magic, type tag or string offsets ML_ERR_INTERNAL. */
ml_status ml_mlb1_read(ml_ctx* ctx, const void* src, int64_t len, int64_t* consumed, ml_result** out);

/* ------------------------------------------------------------------ */
/* Stateless operators                                                */
/* ------------------------------------------------------------------ */

/* Result type of `expr` over `batch`'s columns (after the coercions of
 * crates/minilake-exec/src/expr/eval.rs:218-301). */
ml_status ml_expr_type(ml_ctx* ctx, const ml_expr* expr, const ml_batch* batch, int32_t* out_type);

/* FilterOperator::execute (crates/minilake-exec/src/operators/filter.rs:19-25)
 * via select (expr/eval.rs:167-196) + mask_to_selection (kernels/cmp.rs:85-118):
 * writes the ascending indices of active rows where `predicate` is TRUE
 * (NULL dropped) to `out_sel` (device, capacity = active rows). */
ml_status ml_filter(ml_ctx* ctx, const ml_expr* predicate, const ml_batch* in,
                    uint32_t* out_sel, int64_t* out_len);

/* ProjectionOperator::execute (crates/minilake-exec/src/operators/projection.rs:21-32)
 * via evaluate_to_column (expr/eval.rs:154-160): evaluates each expression over
 * every physical row; the selection is kept by the caller. Result column i holds
 * exprs[i] (with its validity bitmap if any row is NULL). */
ml_status ml_project(ml_ctx* ctx, const ml_expr* exprs, int32_t num_exprs, const ml_batch* in,
                     ml_result** out);

/* hash_columns (crates/minilake-exec/src/kernels/hash.rs:16-23): one 64-bit
 * hash per active row over the key expressions, bit-identical to the reference
 * (SEED start, combine per column, hash_bytes per dictionary entry). */
ml_status ml_hash_columns(ml_ctx* ctx, const ml_expr* keys, int32_t num_keys, const ml_batch* in,
                          uint64_t* out_hashes);

/* ------------------------------------------------------------------ */
/* Hash aggregate sink (HashAggregateSink / HashAggState,
 * crates/minilake-exec/src/operators/aggregate/hash.rs:55-172,285-484;
 * accumulators aggregate/accumulator.rs:99-544)                      */
/* ------------------------------------------------------------------ */
enum { /* AggregateFunction, accumulator.rs:24-38 */
  ML_AGG_COUNT_STAR = 0, ML_AGG_COUNT = 1, ML_AGG_SUM = 2, ML_AGG_AVG = 3,
  ML_AGG_MIN = 4, ML_AGG_MAX = 5
};
enum { /* AggMode, aggregate/mod.rs:50-57 */
  ML_AGG_MODE_SINGLE = 0, ML_AGG_MODE_PARTIAL = 1, ML_AGG_MODE_FINAL = 2
};
enum { /* how the group table is sized */
  ML_SIZING_PLAIN = 0,   /* start small, double and rehash before a batch that could
                            exceed 7/8 load (swiss.rs:219-232 applied per device batch) */
  ML_SIZING_HINT = 1,    /* pre-size from capacity_hint (room for at least that many
                            groups); more groups than that room -> ML_ERR_CAPACITY */
  ML_SIZING_LEARNED = 2  /* pre-size from the learned estimator; more groups than the
                            room it gave -> a partitioned table is partitioned again at
                            the size its filled partitions project (at most the largest
                            partitioned table), a global one falls back to
                            ML_SIZING_PLAIN; both automatically, with identical results */
};

typedef struct ml_aggregate {
  int32_t func;     /* ML_AGG_* */
  int32_t arg_type; /* type of the argument expression (ignored for COUNT(*)) */
  ml_expr arg;      /* empty for COUNT(*) */
} ml_aggregate;

typedef struct ml_agg_spec {
  const ml_expr* group_exprs;
  const int32_t* key_types;
  int32_t num_keys;
  int32_t num_aggs;
  const ml_aggregate* aggs;
  int32_t mode;          /* ML_AGG_MODE_SINGLE only on the GPU */
  int32_t sizing;        /* ML_SIZING_* */
  int64_t capacity_hint; /* expected distinct groups (HINT / LEARNED) */
} ml_agg_spec;

/* With num_keys == 0 the sink is UngroupedAggregateSink
 * (aggregate/ungrouped.rs:36-139): one output row, also over no input rows
 * (COUNT 0, SUM/AVG/MIN/MAX NULL). */
ml_status ml_hashagg_create(ml_ctx* ctx, const ml_agg_spec* spec, ml_hashagg** out);
/* HashAggLocal::sink (aggregate/hash.rs:429-437). */
ml_status ml_hashagg_sink(ml_hashagg* agg, const ml_batch* batch);
/* HashAggregateSink::finalize (aggregate/hash.rs:369-425): key columns then one
 * column per aggregate (finish semantics of accumulator.rs:500-526). */
ml_status ml_hashagg_finalize(ml_hashagg* agg, ml_result** out);
/* Groups currently held (valid at any point; synchronizes). */
ml_status ml_hashagg_num_groups(ml_hashagg* agg, int64_t* out);
/* Slots allocated and number of grow-and-rehash steps taken so far. */
ml_status ml_hashagg_table_stats(ml_hashagg* agg, int64_t* capacity, int32_t* resizes);
void ml_hashagg_free(ml_hashagg* agg);

/* ------------------------------------------------------------------ */
/* Hash join (crates/minilake-exec/src/operators/join.rs:36-338)      */
/* ------------------------------------------------------------------ */
typedef struct ml_join_build_spec {
  const ml_expr* keys;       /* build key expressions */
  const int32_t* key_types;  /* shared key types (physical_planner.rs:112-121 casts) */
  int32_t num_keys;
  int32_t _pad;
} ml_join_build_spec;

/* JoinBuildSink::new / create_local / sink (join.rs:72-101,217-222). */
ml_status ml_join_build_create(ml_ctx* ctx, const ml_join_build_spec* spec, ml_join_build** out);
ml_status ml_join_build_sink(ml_join_build* b, const ml_batch* batch);
/* JoinBuildSink::finalize + build_table (join.rs:122-203). The table keeps every
 * build column; NULL keys never match. Consumes the builder. */
ml_status ml_join_build_finalize(ml_join_build* b, ml_join_table** out);
/* Finalize n builds of one context together (each as ml_join_build_finalize;
 * every builder is consumed). Their layout decisions share one read of the
 * device, so independent build sides cost one host round trip. On failure
 * every out[i] is NULL. */
ml_status ml_join_build_finalize_many(ml_join_build** builds, int32_t n, ml_join_table** out);
int64_t ml_join_table_num_rows(const ml_join_table* t);
/* 1 if every build key is distinct (each probe row matches at most once). */
int32_t ml_join_table_unique(const ml_join_table* t);
/* The build side's columns in table row order (JoinTable.batch, join.rs:36-45);
 * build row indices returned by ml_join_probe index these. Row order is
 * internal to the table: one unique integer key over a small range is stored
 * in key order (rows with a NULL key last). ml_join_table_column waits for the
 * context's stream first (the build's last kernels may still be queued), so
 * the columns can then be read from any stream. */
int32_t ml_join_table_num_columns(const ml_join_table* t);
const ml_column* ml_join_table_column(const ml_join_table* t, int32_t i);
void ml_join_table_free(ml_join_table* t);

/* HashJoinProbe::execute (join.rs:263-315): all matching (probe row, build row)
 * pairs. Result columns: probe row index (INT64, physical row of `probe`) and
 * build row index (INT64). Pair order is unspecified. */
ml_status ml_join_probe(ml_ctx* ctx, const ml_join_table* t, const ml_expr* probe_keys,
                        int32_t num_keys, const ml_batch* probe, ml_result** out_pairs);

/* ------------------------------------------------------------------ */
/* Fused pipelines: Source -> Operator* -> Sink
 * (Pipeline::run_worker/push, crates/minilake-exec/src/pipeline.rs:109-161).
 * One kernel runs the whole chain per row; nothing between operators is
 * materialized in HBM.                                               */
/* ------------------------------------------------------------------ */
enum { ML_PIPE_FILTER = 0, ML_PIPE_PROBE = 1, ML_PIPE_PROJECT = 2 };

typedef struct ml_operator {
  int32_t kind;
  int32_t build_is_left;      /* PROBE: output layout build ++ probe (join.rs:331-335) */
  ml_expr predicate;          /* FILTER */
  const ml_join_table* table; /* PROBE */
  const ml_expr* keys;        /* PROBE: probe key expressions */
  int32_t num_keys;           /* PROBE */
  int32_t num_exprs;          /* PROJECT */
  const ml_expr* exprs;       /* PROJECT: the new layout is these expressions */
} ml_operator;

/* Run a pipeline whose sink is a hash aggregate / join build / TopN. */
ml_status ml_pipeline_hashagg(ml_ctx* ctx, const ml_batch* source, const ml_operator* ops,
                              int32_t num_ops, ml_hashagg* sink);
ml_status ml_pipeline_join_build(ml_ctx* ctx, const ml_batch* source, const ml_operator* ops,
                                 int32_t num_ops, ml_join_build* sink);
ml_status ml_pipeline_topn(ml_ctx* ctx, const ml_batch* source, const ml_operator* ops,
                           int32_t num_ops, ml_topn* sink);
/* Pipeline ending in a collect sink (CollectSink, operators/collect.rs:11-53):
 * every column of the final layout, compacted to the surviving rows. When a
 * probe can match several build rows the output size is unknown, so an exact
 * counting pass runs first. */
ml_status ml_pipeline_collect(ml_ctx* ctx, const ml_batch* source, const ml_operator* ops,
                              int32_t num_ops, ml_result** out);
/* LimitSink (operators/limit.rs:26-117): rows [offset, offset + limit) of the
 * pipeline's output, in source order (the reference keeps whichever rows
 * arrive first; without ORDER BY any such rows are a correct answer). */
ml_status ml_pipeline_limit(ml_ctx* ctx, const ml_batch* source, const ml_operator* ops, int32_t num_ops,
                            int64_t limit, int64_t offset, ml_result** out);
/* The same with the output pre-sized to `capacity_hint` rows (a learned
 * estimate) instead of the counting pass; if the rows do not fit, the exact
 * path reruns (`*used_fallback` = 1). Results are identical either way. */
ml_status ml_pipeline_collect_sized(ml_ctx* ctx, const ml_batch* source, const ml_operator* ops,
                                    int32_t num_ops, int64_t capacity_hint, int32_t* used_fallback,
                                    ml_result** out);
/* Rows reaching the end of the pipeline (COUNT(*) over a join), no output. */
ml_status ml_pipeline_count(ml_ctx* ctx, const ml_batch* source, const ml_operator* ops, int32_t num_ops,
                            int64_t* out_rows);

/* ------------------------------------------------------------------ */
/* TopN sink (SortSink with limit, crates/minilake-exec/src/operators/sort.rs:43-268) */
/* ------------------------------------------------------------------ */
typedef struct ml_sort_key {
  ml_expr expr;
  int32_t asc;
  int32_t nulls_first;
} ml_sort_key;

typedef struct ml_sort_spec {
  const ml_sort_key* keys;
  int32_t num_keys;
  int32_t _pad;
  int64_t limit; /* ORDER BY ... LIMIT n (physical_planner.rs:150-178); < 0: ORDER BY alone (a full sort) */
} ml_sort_spec;

ml_status ml_topn_create(ml_ctx* ctx, const ml_sort_spec* spec, ml_topn** out);
ml_status ml_topn_sink(ml_topn* t, const ml_batch* batch);
/* Best `limit` rows, sorted; every column of the sunk batches is gathered. */
ml_status ml_topn_finalize(ml_topn* t, ml_result** out);
void ml_topn_free(ml_topn* t);

/* ------------------------------------------------------------------ */
/* Per-thread sink states: LocalSink (crates/minilake-exec/src/pipeline.rs:70-79),
 * as MiniLake's run_parallel drives them (scheduler.rs:102-144): each worker
 * thread makes a local state on its own context with create_local, sinks its
 * morsels into it (ml_*_sink / ml_pipeline_* with the local handle) without
 * locking, and calls combine exactly once; combine merges the local into the
 * global sink under the global's lock (from any thread) and frees the local.
 * finalize on the global runs once, after every combine. A global's context
 * is used by combine and finalize: its owner does not use it meanwhile.     */
/* ------------------------------------------------------------------ */
/* HashAggLocal::combine (aggregate/hash.rs:439-483): the local's groups as
 * partial states, merged by the global's Final-mode merge. The global sink
 * may also take batches itself; finalize then merges its own groups too. */
ml_status ml_hashagg_create_local(ml_hashagg* global, ml_ctx* ctx, ml_hashagg** out);
ml_status ml_hashagg_combine(ml_hashagg* global, ml_hashagg* local);
/* JoinBuildLocal::combine (join.rs:205-226): the local's build rows are
 * appended to the global's (the string dictionaries must be the same). */
ml_status ml_join_build_create_local(ml_join_build* global, ml_ctx* ctx, ml_join_build** out);
ml_status ml_join_build_combine(ml_join_build* global, ml_join_build* local);
/* SortLocal::combine (sort.rs:253-268): the local's best rows merged into
 * the global's best (a full sort merges everything). */
ml_status ml_topn_create_local(ml_topn* global, ml_ctx* ctx, ml_topn** out);
ml_status ml_topn_combine(ml_topn* global, ml_topn* local);

/* LimitSink (operators/limit.rs:26-117). Its local state holds nothing, so
 * the one handle is shared: ml_limit_sink may be called from several threads
 * at once, each with its own context (the batch is compacted there); rows are
 * kept in arrival order until offset + limit are held, after which
 * ml_limit_is_finished (LocalSink::is_finished, limit.rs:114-116) is 1 and
 * further batches are ignored. finalize returns rows [offset, offset + limit)
 * of what was kept (no columns if no batch arrived). */
typedef struct ml_limit ml_limit;
ml_status ml_limit_create(ml_ctx* ctx, int64_t limit, int64_t offset, ml_limit** out);
ml_status ml_limit_sink(ml_limit* s, ml_ctx* ctx, const ml_batch* batch);
int32_t ml_limit_is_finished(const ml_limit* s);
ml_status ml_limit_finalize(ml_limit* s, ml_result** out);
void ml_limit_free(ml_limit* s);

/* ------------------------------------------------------------------ */
/* Batched fragments: many independent
 *   SELECT g, SUM(v), COUNT(*) FROM probe JOIN build USING (key) GROUP BY g
 * fragments (g = the group id of the matching build row) in one persistent-
 * kernel launch sequence, optionally replayed through a captured hipGraph.
 * A probe row adds to the group of every build row with an equal key (the
 * reference's HashJoinProbe emits every chain match, join.rs:283-300); NULL
 * keys never match. No reference counterpart as a whole: the nearest analog
 * is the morsel queue, crates/minilake-exec/src/scheduler.rs:44-68.        */
/* ------------------------------------------------------------------ */
typedef struct ml_fragment {
  const int64_t* build_keys;       /* device */
  const int64_t* build_groups;     /* device, group id of each build row; must lie in [0, num_groups) */
  const uint64_t* build_validity;  /* device LSB-first bitmap of the build keys; NULL = all valid */
  int64_t build_rows;
  const int64_t* probe_keys;       /* device */
  const double* probe_values;      /* device */
  const uint64_t* probe_validity;  /* device LSB-first bitmap of the probe keys; NULL = all valid */
  int64_t probe_rows;
  int64_t num_groups;              /* group id domain, >= 1 */
  double* out_sums;                /* device, num_groups; overwritten by every run */
  int64_t* out_counts;             /* device, num_groups; matched pairs per group */
} ml_fragment;

typedef struct ml_fragment_tuning {
  int32_t block_size;      /* threads per workgroup: 256, 512 or 1024 */
  int32_t tile_rows;       /* rows per work item */
  int32_t blocks_per_cu;   /* persistent grid = CUs * blocks_per_cu */
  int32_t lds_groups;      /* LDS aggregation entries (a power of two <= 2048; 0 = none): domains up
                              to this size aggregate in an LDS array */
  int32_t probe_width;     /* probe rows each lane keeps in flight: 1, 2, 4 or 8 */
  int32_t lds_cache;       /* 1: larger domains aggregate through an LDS cache of lds_groups entries */
} ml_fragment_tuning;

typedef struct ml_fragments ml_fragments;

/* Plan a set of fragments: allocates every hash table and the work queues.
 * Sizes and null pointers are checked here (ML_ERR_PLAN). */
ml_status ml_fragments_create(ml_ctx* ctx, const ml_fragment* frags, int32_t n,
                              const ml_fragment_tuning* tuning, ml_fragments** out);
/* Run all fragments (reset, build and probe launches over device work
 * queues) and wait for them. A build row whose group id lies outside
 * [0, num_groups) is skipped on the device and the call returns
 * ML_ERR_PLAN naming the first such fragment. */
ml_status ml_fragments_run(ml_fragments* f);
/* The same launch sequence captured into a hipGraph (once) and replayed. */
ml_status ml_fragments_run_graph(ml_fragments* f);
/* Baseline: the same kernels launched fragment by fragment. */
ml_status ml_fragments_run_per_fragment(ml_fragments* f);
/* Autotune: coordinate descent over probe width, tile rows, block size,
 * blocks per CU and the LDS aggregation mode, timing whole runs; keeps and
 * returns the fastest tuning. */
ml_status ml_fragments_autotune(ml_fragments* f, ml_fragment_tuning* best);
void ml_fragments_free(ml_fragments* f);

/* ------------------------------------------------------------------ */
/* Learned cardinality estimation (replaces the crude estimate_rows,
 * crates/minilake-sql/src/physical_planner.rs:251-295, for table sizing).
 * MLP 32 -> 64 -> 64 -> 2, FP32, ReLU.                                 */
/* ------------------------------------------------------------------ */
#define ML_MLP_IN 32
#define ML_MLP_H1 64
#define ML_MLP_H2 64
#define ML_MLP_OUT 2
#define ML_MLP_NUM_PARAMS \
  (ML_MLP_IN * ML_MLP_H1 + ML_MLP_H1 + ML_MLP_H1 * ML_MLP_H2 + ML_MLP_H2 + ML_MLP_H2 * ML_MLP_OUT + ML_MLP_OUT)

typedef struct ml_mlp ml_mlp;

/* Parameters in host memory: W1[64][32], b1[64], W2[64][64], b2[64], W3[2][64], b3[2]. */
ml_status ml_mlp_create(ml_ctx* ctx, const float* params, ml_mlp** out);
ml_status ml_mlp_get_params(ml_mlp* m, float* params);
/* Forward on device features x[n][32] -> y[n][2]. */
ml_status ml_mlp_forward(ml_mlp* m, const float* x, int64_t n, float* y);
/* Mean-squared-error loss and parameter gradients on one batch (device x, t). */
ml_status ml_mlp_loss_grad(ml_mlp* m, const float* x, const float* t, int64_t n, float* loss,
                           float* grads_host);
/* One Adam step over a minibatch; returns the batch loss. */
ml_status ml_mlp_train_step(ml_mlp* m, const float* x, const float* t, int64_t n, float lr,
                            float* loss);
void ml_mlp_free(ml_mlp* m);

/* Features of one aggregation/join input (device key column): row-group
 * min/max/null-count statistics (crates/minilake-storage/src/stats.rs:15-22)
 * plus a KMV sketch of a `sample`-row sample, 32 floats. */
ml_status ml_sizing_features(ml_ctx* ctx, const int64_t* keys, int64_t n, int64_t row_group_rows,
                             int32_t sample, uint64_t seed, float* out_features);
/* Attach a trained estimator to the context: ML_SIZING_LEARNED sinks use it. */
ml_status ml_ctx_set_estimator(ml_ctx* ctx, ml_mlp* m);

#ifdef __cplusplus
}
#endif
#endif /* MINILAKE_GPU_H */
