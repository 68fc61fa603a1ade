// Host-side structures of the gfx950 execution path (C++, no exceptions cross
// the C ABI: every entry point catches and returns an ml_status).
#pragma once

#include <hip/hip_runtime.h>

#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/minilake_gpu.h"
#include "kernels.h"
#include "mlgpu.h"

namespace mlg {

// MiniLakeError as a C++ exception inside the library.
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
[[noreturn]] inline void fail(int code, const std::string& m) { throw Error(code, m); }

inline void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) fail(ML_ERR_INTERNAL, std::string(what) + ": " + hipGetErrorString(e));
}

inline const char* type_name(int t) {
  switch (t) {
    case ML_TYPE_BOOLEAN: return "BOOLEAN";
    case ML_TYPE_INT32: return "INT32";
    case ML_TYPE_INT64: return "INT64";
    case ML_TYPE_FLOAT64: return "FLOAT64";
    case ML_TYPE_DATE: return "DATE";
    case ML_TYPE_UTF8: return "VARCHAR";
    default: return "NULL";
  }
}
inline int type_width(int t) {
  switch (t) {
    case ML_TYPE_BOOLEAN: return 1;
    case ML_TYPE_INT32: case ML_TYPE_DATE: case ML_TYPE_UTF8: return 4;
    default: return 8;
  }
}

// ------------------------------------------------------------------ memory
// Stream-ordered caching allocator: blocks are recycled in power-of-two size
// classes and only returned to the driver when the context is destroyed. All
// work of a context runs on its one stream, so reuse after free is ordered.
class DeviceArena {
 public:
  explicit DeviceArena(int device) : device_(device) {}
  ~DeviceArena();
  void* alloc(size_t bytes);
  void free(void* p);
  size_t bytes_in_use() const { return in_use_; }

 private:
  int device_;
  std::map<size_t, std::vector<void*>> free_;
  std::map<void*, size_t> live_;
  size_t in_use_ = 0;
};

// A device allocation; it keeps its arena alive, so results and tables may
// outlive the context that made them.
struct Buffer {
  std::shared_ptr<DeviceArena> arena;
  void* ptr = nullptr;
  size_t bytes = 0;
  Buffer() = default;
  Buffer(const std::shared_ptr<DeviceArena>& a, size_t n) : arena(a), ptr(n ? a->alloc(n) : nullptr), bytes(n) {}
  Buffer(const Buffer&) = delete;
  Buffer& operator=(const Buffer&) = delete;
  Buffer(Buffer&& o) noexcept { *this = std::move(o); }
  Buffer& operator=(Buffer&& o) noexcept {
    if (this != &o) {
      reset();
      arena = std::move(o.arena); ptr = o.ptr; bytes = o.bytes;
      o.ptr = nullptr; o.bytes = 0;
    }
    return *this;
  }
  ~Buffer() { reset(); }
  void reset() {
    if (ptr && arena) arena->free(ptr);
    ptr = nullptr;
    bytes = 0;
    arena.reset();
  }
  template <class T> T* as() const { return static_cast<T*>(ptr); }
};

// --------------------------------------------------------------- dictionary
// Host copy of an ml_dict, shared by columns and results.
struct HostDict {
  std::vector<std::string> entries;
  std::vector<uint32_t> offsets;
  std::string bytes;
  ml_dict view{};
  static std::shared_ptr<HostDict> from(const ml_dict* d);
  void finish_view();
};

// Device memory handed out front to back from one fixed region (the
// allocation-free entry points carve every buffer of a call from it instead
// of the context's pool). Offsets are 256-byte aligned.
struct ScratchBump {
  uint8_t* base = nullptr;
  size_t cap = 0;
  size_t used = 0;
  void* take(size_t n) {
    const size_t at = (used + 255) & ~(size_t)255;
    if (at + n > cap)
      fail(ML_ERR_RESOURCES_EXHAUSTED, "scratch region of " + std::to_string(cap) + " bytes exhausted (" +
                                           std::to_string(at + n) + " bytes needed)");
    used = at + n;
    return base + at;
  }
};

// ------------------------------------------------------------------ context
struct MlpModel;

// MemoryPool (memory.rs:22-76): a budget for pipeline-breaker state (group
// tables, join tables); limit 0 = unlimited. With spill on, a hash aggregate
// over budget writes its partial states to host memory.
struct MemPool {
  int64_t limit = 0;
  int64_t used = 0;
  int64_t peak = 0;
  bool spill = false;
  bool try_grow(int64_t bytes) {
    if (bytes > 0 && limit > 0 && used + bytes > limit) return false;
    used += bytes;
    if (used > peak) peak = used;
    return true;
  }
  void shrink(int64_t bytes) { used -= bytes < used ? bytes : used; }
};

// Pinned host words that receive the flags of checks left running on the
// stream (a join build's duplicate-key or partition-overflow test), so the
// host need not wait for them when the build is launched. Shared by a
// context and the tables it built.
struct CheckPool {
  static constexpr int WORDS = 1024;
  uint32_t* words = nullptr;
  std::vector<int> free;
  // events recorded right after each word's copy: a table's check is read by
  // waiting for its own copy, not for everything the stream holds after it
  std::vector<hipEvent_t> events;
  CheckPool();
  ~CheckPool();
  int take();  // -1 when every word is in use
  void give(int w) { if (w >= 0) free.push_back(w); }
  hipEvent_t event(int w);
};
}  // namespace mlg

struct ml_ctx {
  ml_exec_config cfg{};
  int device = 0;
  int num_cus = 256;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  std::string last_error;
  std::shared_ptr<mlg::DeviceArena> arena;
  mlg::MlpModel* estimator = nullptr;
  // small per-context device scratch: counters and flags
  mlg::Buffer counters;  // 64 x u64
  static constexpr int DONE_WORD = 63;  // counters word: workgroups finished (kept at 0 between launches)
  // pinned host staging for small device-to-host reads (counts, flags), so
  // they skip the pageable-memory staging path of hipMemcpy
  void* pinned = nullptr;
  static constexpr size_t PINNED_BYTES = 4096;
  std::shared_ptr<mlg::CheckPool> checks;
  // pinned ring for small uploads (instruction arrays, small result columns):
  // a region is rewritten only after the ring wraps, which first waits for
  // the stream, so no upload has to wait for its copy
  uint8_t* ring = nullptr;
  size_t ring_head = 0;
  static constexpr size_t RING_BYTES = 1 << 20;
  uint64_t check_seq = 0;   // deferred checks enqueued so far
  uint64_t synced_seq = 0;  // deferred checks the stream has completed (seen at a synchronization)
  std::map<const ml_dict*, std::shared_ptr<mlg::HostDict>> dicts;
  std::shared_ptr<mlg::HostDict> dict(const ml_dict* d);
  // the budget for pipeline-breaker state (shared with the join tables built here)
  std::shared_ptr<mlg::MemPool> mem = std::make_shared<mlg::MemPool>();
  // launch statistics
  bool stats_on = false;
  int64_t stats_min_rows = 0;  // record launches over at least this many source rows
  struct PendingStat {
    int sink;
    int64_t rows;
    hipEvent_t start, stop;
  };
  std::vector<PendingStat> stats;
  std::vector<hipEvent_t> event_pool;  // events of taken stats, reused
  hipEvent_t pooled_event();
  void launch(const mlg::Plan& P, int grid, int block, size_t lds);
  // Immutable device copies of small host arrays (instruction arrays, lookup
  // tables), keyed by their bytes: a plan compiled again from the same
  // expressions over the same columns reuses them instead of copying again.
  // Holders keep an entry alive after the cache drops it.
  std::shared_ptr<mlg::Buffer> const_upload(const void* p, size_t n);
  std::unordered_map<std::string, std::shared_ptr<mlg::Buffer>> const_cache;
  size_t const_cache_bytes = 0;
  static constexpr size_t CONST_CACHE_BYTES = 4 << 20;
  static constexpr size_t CONST_CACHE_ENTRIES = 4096;
};

namespace mlg {

// Enqueue a host-to-device copy of n bytes through the context's pinned
// upload ring: returns without waiting, and `src` may be freed at once.
void ring_upload(ml_ctx* ctx, void* dst, const void* src, size_t n);

// ------------------------------------------------------------ result column
struct ResultColumn {
  int type = ML_TYPE_INT64;
  Buffer data;
  Buffer valid;                     // u8 per row, or empty
  std::shared_ptr<HostDict> dict;   // UTF8
  Buffer bitmap;                    // LSB-first validity words exported via ml_column
};

}  // namespace mlg

struct ml_result {
  ml_ctx* ctx = nullptr;
  int device = 0;
  int64_t num_rows = 0;
  std::vector<mlg::ResultColumn> cols;
  std::vector<ml_column> views;
  void finish_views();
};

namespace mlg {

// -------------------------------------------------------------- expressions
struct Scalar {
  int type = ML_TYPE_NULL;  // ML_TYPE_NULL = SQL NULL
  int64_t i = 0;
  double f = 0;
  std::string s;
  bool is_null() const { return type == ML_TYPE_NULL; }
};

// A column of a pipeline's row layout.
struct VCol;
struct Node;
using NodePtr = std::shared_ptr<const Node>;

struct VCol {
  int type = ML_TYPE_INT64;
  int stage = 0;
  const void* data = nullptr;       // device
  const uint8_t* valid = nullptr;   // device u8 validity or null
  std::shared_ptr<HostDict> dict;   // UTF8
  NodePtr expr;                     // a projected expression (inlined where used)
};

struct Node {
  int kind = ML_EXPR_COLUMN;
  int op = 0;
  int type = 0;
  int column = 0;
  Scalar lit;
  NodePtr left, right;
  const VCol* col = nullptr;  // resolved physical column
};

NodePtr parse_expr(const ml_expr& e);

// Compiles expression trees against a row layout into device instructions.
class Compiler {
 public:
  Compiler(ml_ctx* ctx, std::vector<Instr>& ins, std::vector<Buffer>& tables)
      : ctx_(ctx), ins_(ins), tables_(tables) {}
  // Result type of a resolved tree (the reference's expression typing).
  int type_of(const NodePtr& n);
  // Compile a resolved tree; the value is cast to `want` (evaluate_to_column
  // semantics) unless want < 0. Returns the program.
  Prog compile(const NodePtr& n, int want, int* out_type = nullptr);
  // Group key word: cast to `key_type`; f64 -0.0 -> +0.0 and each NaN a word
  // of its own, numbered from the device counter `nan_ids` (null: one
  // canonical NaN, for programs that are only inspected, never run).
  Prog compile_key(const NodePtr& n, int key_type, uint64_t* nan_ids = nullptr);
  // Join key word: cast to `key_type`; f64 -0.0 -> +0.0 and NaN -> NULL.
  Prog compile_join_key(const NodePtr& n, int key_type);
  // hash_columns input: a word to hash, or an already-hashed value (dictionaries).
  Prog compile_hash_input(const NodePtr& n, bool* prehashed);
  // A plain load of a layout column.
  Prog compile_load(const VCol& c);
  // When set, lookup tables are carved from this region, not the context's pool.
  ScratchBump* bump = nullptr;
  // lookup tables shared through the context's constant cache
  std::vector<std::shared_ptr<Buffer>> consts;

 private:
  struct Val {
    bool is_const;
    Scalar s;   // constant
    int type;   // runtime type when not constant
  };
  Val gen(const NodePtr& n);               // emits code for non-constant values
  Val gen_binary(const Node& n);
  Val materialize(const Val& v);           // constant -> CONST instruction
  void emit_cast(int from, int to);
  Val cast_val(const Val& v, int to);
  void push(uint16_t op, uint64_t imm = 0, uint64_t imm2 = 0, uint8_t stage = 0, uint32_t flags = 0);
  void unary(uint16_t op, uint64_t imm = 0);
  void binary(uint16_t op);
  void fuse(int32_t first);
  const uint8_t* dict_predicate(const HostDict& d, int op, const std::string& lit);

  ml_ctx* ctx_;
  std::vector<Instr>& ins_;
  std::vector<Buffer>& tables_;
  int depth_ = 0;
  int max_depth_ = 0;
  bool nullable_ = false;
  std::vector<const Node*> consts_;
};

// ------------------------------------------------------------- join tables
struct JoinTableImpl {
  int64_t nrows = 0;
  std::vector<int> key_types;
  std::vector<std::shared_ptr<HostDict>> key_dicts;
  std::vector<int> col_types;
  std::vector<std::shared_ptr<HostDict>> col_dicts;
  std::vector<Buffer> cols;
  std::vector<Buffer> col_valid;
  Buffer slots;
  Buffer bloom;
  Buffer dense_bits, dense_rank, dense_keys;  // dense layout (DevJoinTable::dense)
  Buffer dense_keys8;  // a byte copy of the second key words, used once they are known to fit
  // Dense, not every key present, rows held at their key offset (no rank
  // directory, dev.dense_rank null): positions 0..dense_range-1 by key, then
  // the NULL-key rows, offset_rows in all. The C ABI numbers build rows in key
  // order; that numbering comes from a rank of the presence bits computed on
  // first use (probe pairs, column views), with compacted column copies.
  bool dense_offset = false;
  int64_t offset_rows = 0;
  // A full or offset-indexed table's build column that is its first key
  // (the same program): not stored, since row p holds key min + p. Probes
  // read the probe key in its place; the views rebuild it on first access.
  int implied_col = -1;
  Buffer offset_rank;
  std::vector<Buffer> compact_cols, compact_valid;
  DevJoinTable dev{};
  std::vector<ml_column> views;
  // A build whose last check still runs on the stream: its flags arrive in
  // checks->words[check]; join_table_resolve finishes the table before its
  // layout is used on the host.
  int pending = 0;  // PENDING_*
  int check = -1;
  uint64_t check_seq = 0;
  std::shared_ptr<CheckPool> checks;
  Buffer keys, key_null;                      // key words, for a fallback build
  std::vector<Buffer> orig_cols, orig_valid;  // dense: rows in build order
};
enum : int { PENDING_NONE = 0, PENDING_DENSE = 1, PENDING_PARTS = 2 };

}  // namespace mlg

struct ml_join_table {
  ml_ctx* ctx = nullptr;
  mlg::JoinTableImpl impl;
  std::shared_ptr<mlg::MemPool> pool;  // the budget its bytes are reserved against
  int64_t reserved = 0;
  ~ml_join_table() {
    if (impl.checks) impl.checks->give(impl.check);
    if (pool) pool->shrink(reserved);
  }
};

namespace mlg {
// Look up this device's address of the allocation-free entry points' scratch
// array (once per device; call with the device current).
void noalloc_setup(int device);
// Finish a table whose build check is pending (waits for the stream once).
void join_table_resolve(ml_join_table* t);
}  // namespace mlg

// -------------------------------------------------------------- the pipeline
namespace mlg {

struct PipelineBuild {
  ml_ctx* ctx;
  Plan P{};
  std::vector<Instr> ins;
  std::vector<Buffer> tables;  // device lookup tables referenced by instructions
  std::vector<const VCol*> layout;
  std::vector<std::unique_ptr<VCol>> owned;
  std::vector<std::vector<uint8_t>> host_tables;
  int nstages = 1;
  int64_t active_rows = 0;
  bool expanding = false;
  Compiler comp;
  std::shared_ptr<Buffer> dev_ins;  // the instructions, from the context's constant cache
  // When set, every device buffer of the plan (expanded validity, lookup
  // tables, instructions) is carved from this region, not the context's pool.
  ScratchBump* bump = nullptr;
  void use_scratch(ScratchBump* b) { bump = b; comp.bump = b; }

  PipelineBuild(ml_ctx* c) : ctx(c), comp(c, ins, tables) {}
  void source(const ml_batch* b, bool use_selection = true);
  void add_ops(const ml_operator* ops, int n);
  void upload();  // instructions -> device; fills P.ins
  int grid(int block = 256) const;
  NodePtr resolve(const ml_expr& e) const;
  const VCol* add_col(VCol c);
};

NodePtr resolve_layout(const NodePtr& n, const std::vector<const VCol*>& layout);

}  // namespace mlg
