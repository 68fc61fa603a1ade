// CPU ORACLE — test infrastructure only (see ref_core.h). extern "C" surface
// used by tests/ and by bench.py's CPU-baseline leg through ctypes.
#pragma once
#include <stdint.h>

#include "../include/minilake_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mlo_options {
  int32_t threads;     // worker threads (ExecConfig::threads)
  int32_t batch_size;  // rows per batch (DEFAULT_BATCH_SIZE = 2048)
  int64_t morsel_rows; // rows per morsel (a row group: 122,880)
} mlo_options;

typedef struct mlo_table mlo_table;
typedef struct mlo_join mlo_join;

const char* mlo_last_error(void);
uint64_t mlo_fold_mul(uint64_t a, uint64_t b);
uint64_t mlo_hash_u64(uint64_t x);
uint64_t mlo_combine(uint64_t h, uint64_t v);
uint64_t mlo_hash_bytes(const uint8_t* b, uint64_t n);
uint64_t mlo_match_tag(uint64_t group, uint8_t tag);
uint64_t mlo_match_empty(uint64_t group);
int64_t mlo_swiss_group_ids(const uint64_t* keys, int64_t n, int64_t batch, int32_t weak, uint32_t* out_ids,
                            int64_t* out_capacity, int32_t* out_resizes);
int mlo_expr_type(const ml_expr* e, const ml_batch* b, int32_t* out);
int mlo_filter(const ml_expr* pred, const ml_batch* in, uint32_t* out_sel, int64_t* out_len);
int mlo_project(const ml_expr* exprs, int32_t n, const ml_batch* in, mlo_table** out);
int mlo_hash_columns(const ml_expr* keys, int32_t nkeys, const ml_batch* in, uint64_t* out);
int mlo_pipeline_hashagg(const ml_batch* src, const ml_operator* ops, int32_t nops, const ml_agg_spec* spec,
                         const mlo_options* opt, mlo_table** out);
int mlo_pipeline_join_build(const ml_batch* src, const ml_operator* ops, int32_t nops,
                            const ml_join_build_spec* spec, const mlo_options* opt, mlo_join** out);
int64_t mlo_join_num_rows(const mlo_join* j);
int64_t mlo_join_num_keys(const mlo_join* j);
void mlo_join_free(mlo_join* j);
int mlo_join_probe(const mlo_join* j, const ml_expr* keys, int32_t nkeys, const ml_batch* probe, mlo_table** out);
int mlo_pipeline_topn(const ml_batch* src, const ml_operator* ops, int32_t nops, const ml_sort_spec* spec,
                      const mlo_options* opt, mlo_table** out);
// LimitSink (operators/limit.rs:26-117): rows [offset, offset + limit) of what the pipeline produces
int mlo_pipeline_limit(const ml_batch* src, const ml_operator* ops, int32_t nops, int64_t limit, int64_t offset,
                       const mlo_options* opt, mlo_table** out);
int mlo_pipeline_collect(const ml_batch* src, const ml_operator* ops, int32_t nops, const mlo_options* opt,
                         mlo_table** out);
int64_t mlo_table_num_rows(const mlo_table* t);
int32_t mlo_table_num_columns(const mlo_table* t);
const ml_column* mlo_table_column(const mlo_table* t, int32_t i);
int32_t mlo_table_resizes(const mlo_table* t);
int64_t mlo_table_capacity(const mlo_table* t);
void mlo_table_free(mlo_table* t);
// MLB1 batch bytes (ipc.rs:44-170): write returns malloc'd bytes (free with
// mlo_bytes_free); read returns *out = NULL at a clean end of stream.
int mlo_mlb1_write(const ml_batch* src, uint8_t** out, int64_t* len);
void mlo_bytes_free(uint8_t* p);
int mlo_mlb1_read(const uint8_t* src, int64_t len, int64_t* consumed, mlo_table** out);

#ifdef __cplusplus
}
#endif
