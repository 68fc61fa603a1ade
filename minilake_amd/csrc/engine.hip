// Host side of the gfx950 execution path: context, device memory, pipeline
// planning and the sinks behind the C ABI (include/minilake_gpu.h).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <numeric>
#include <string_view>
#include <unordered_map>
#include <unordered_set>

#include "engine.h"

namespace mlg {

// ------------------------------------------------------------------ memory
static size_t size_class(size_t n) {
  size_t c = 256;
  while (c < n) c <<= 1;
  return c;
}

void* DeviceArena::alloc(size_t bytes) {
  const size_t c = size_class(bytes);
  auto& fl = free_[c];
  void* p = nullptr;
  if (!fl.empty()) {
    p = fl.back();
    fl.pop_back();
  } else {
    hipError_t e = hipMalloc(&p, c);
    if (e != hipSuccess) {
      // release every cached block and retry once
      (void)hipGetLastError();
      hip_check(hipDeviceSynchronize(), "synchronize before trim");
      for (auto& kv : free_)
        for (void* q : kv.second) (void)hipFree(q);
      free_.clear();
      e = hipMalloc(&p, c);
      if (e != hipSuccess) {
        (void)hipGetLastError();
        fail(ML_ERR_RESOURCES_EXHAUSTED, "device allocation of " + std::to_string(c) + " bytes failed");
      }
    }
  }
  live_[p] = c;
  in_use_ += c;
  return p;
}

void DeviceArena::free(void* p) {
  auto it = live_.find(p);
  if (it == live_.end()) return;
  in_use_ -= it->second;
  free_[it->second].push_back(p);
  live_.erase(it);
}

DeviceArena::~DeviceArena() {
  (void)hipSetDevice(device_);
  for (auto& kv : free_)
    for (void* q : kv.second) (void)hipFree(q);
  for (auto& kv : live_) (void)hipFree(kv.first);
}

CheckPool::CheckPool() {
  hip_check(hipHostMalloc((void**)&words, WORDS * sizeof(uint32_t), hipHostMallocDefault), "hipHostMalloc");
  for (int w = WORDS - 1; w >= 0; w--) free.push_back(w);
}
CheckPool::~CheckPool() {
  for (hipEvent_t e : events)
    if (e) (void)hipEventDestroy(e);
  if (words) (void)hipHostFree(words);
}
hipEvent_t CheckPool::event(int w) {
  if (events.empty()) events.assign(WORDS, nullptr);
  if (!events[(size_t)w]) hip_check(hipEventCreateWithFlags(&events[(size_t)w], hipEventDisableTiming), "hipEventCreate");
  return events[(size_t)w];
}
int CheckPool::take() {
  if (free.empty()) return -1;
  const int w = free.back();
  free.pop_back();
  return w;
}

static void h2d(ml_ctx* ctx, void* dst, const void* src, size_t n) {
  if (n) hip_check(hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, ctx->stream), "copy to device");
}
// Upload through the pinned ring (returns at once; `src` may be freed after
// the call). Larger uploads copy and then wait.
#ifdef ML_COPY_DEBUG
#define ML_CALLER , const char* caller = __builtin_FUNCTION(), int line = __builtin_LINE()
#define ML_NOTE(kind) fprintf(stderr, "[copy] %s %zu %s:%d\n", kind, n, caller, line)
#else
#define ML_CALLER
#define ML_NOTE(kind)
#endif
static void upload(ml_ctx* ctx, void* dst, const void* src, size_t n ML_CALLER) {
  ML_NOTE("h2d");
  if (!n) return;
  if (!ctx->ring || n > ml_ctx::RING_BYTES / 4) {
    h2d(ctx, dst, src, n);
    hip_check(hipStreamSynchronize(ctx->stream), "synchronize");
    return;
  }
  if (ctx->ring_head + n > ml_ctx::RING_BYTES) {
    hip_check(hipStreamSynchronize(ctx->stream), "synchronize");  // every earlier upload has landed
    ctx->ring_head = 0;
  }
  uint8_t* at = ctx->ring + ctx->ring_head;
  std::memcpy(at, src, n);
  hip_check(hipMemcpyAsync(dst, at, n, hipMemcpyHostToDevice, ctx->stream), "copy to device");
  ctx->ring_head += (n + 255) & ~(size_t)255;
}
void ring_upload(ml_ctx* ctx, void* dst, const void* src, size_t n) { upload(ctx, dst, src, n); }
}  // namespace mlg

std::shared_ptr<mlg::Buffer> ml_ctx::const_upload(const void* p, size_t n) {
  std::string key(static_cast<const char*>(p), n);
  auto it = const_cache.find(key);
  if (it != const_cache.end()) return it->second;
  if (const_cache_bytes + n > CONST_CACHE_BYTES || const_cache.size() >= CONST_CACHE_ENTRIES) {
    const_cache.clear();
    const_cache_bytes = 0;
  }
  auto b = std::make_shared<mlg::Buffer>(arena, std::max<size_t>(n, 1));
  mlg::upload(this, b->ptr, p, n);
  const_cache_bytes += n;
  const_cache.emplace(std::move(key), b);
  return b;
}

namespace mlg {
static void d2h(ml_ctx* ctx, void* dst, const void* src, size_t n ML_CALLER) {
  ML_NOTE("d2h");
  const uint64_t seq = ctx->check_seq;
  if (n && ctx->pinned && n <= ml_ctx::PINNED_BYTES) {
    hip_check(hipMemcpyAsync(ctx->pinned, src, n, hipMemcpyDeviceToHost, ctx->stream), "copy to host");
    hip_check(hipStreamSynchronize(ctx->stream), "synchronize");
    std::memcpy(dst, ctx->pinned, n);
  } else {
    if (n) hip_check(hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToHost, ctx->stream), "copy to host");
    hip_check(hipStreamSynchronize(ctx->stream), "synchronize");
  }
  ctx->synced_seq = seq;  // every deferred check enqueued before this copy has landed
}

// Enqueue a copy of the 4-byte flag word `src` into a pinned check word of
// T, to be read by join_table_resolve. False when no check word is free
// (the caller then reads the flag synchronously).
static bool defer_check(ml_ctx* ctx, JoinTableImpl& T, const uint32_t* src, int kind) {
  const int w = ctx->checks ? ctx->checks->take() : -1;
  if (w < 0) return false;
  hip_check(hipMemcpyAsync(ctx->checks->words + w, src, 4, hipMemcpyDeviceToHost, ctx->stream), "copy to host");
  hip_check(hipEventRecord(ctx->checks->event(w), ctx->stream), "record");
  T.checks = ctx->checks;
  T.check = w;
  T.check_seq = ++ctx->check_seq;
  T.pending = kind;
  return true;
}
static void d2d(ml_ctx* ctx, void* dst, const void* src, size_t n ML_CALLER) {
  ML_NOTE("d2d");
  if (n) hip_check(hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToDevice, ctx->stream), "device copy");
}
static void zero(ml_ctx* ctx, void* p, size_t bytes ML_CALLER) {  // a fill kernel, so the work stays capturable
#ifdef ML_COPY_DEBUG
  const size_t n = bytes;
  ML_NOTE("fill");
#endif
  launch_fill_u64((uint64_t*)p, 0, (int64_t)((bytes + 7) / 8), ctx->stream);
}

// ------------------------------------------------------------ dictionaries
std::shared_ptr<HostDict> HostDict::from(const ml_dict* d) {
  auto h = std::make_shared<HostDict>();
  for (int i = 0; i < d->len; i++)
    h->entries.emplace_back((const char*)d->bytes + d->offsets[i], d->offsets[i + 1] - d->offsets[i]);
  h->finish_view();
  return h;
}
void HostDict::finish_view() {
  offsets.assign(1, 0);
  bytes.clear();
  for (auto& e : entries) {
    bytes += e;
    offsets.push_back((uint32_t)bytes.size());
  }
  view.len = (int32_t)entries.size();
  view.offsets = offsets.data();
  view.bytes = (const uint8_t*)bytes.data();
}

static bool same_dict(const std::shared_ptr<HostDict>& a, const std::shared_ptr<HostDict>& b) {
  return a == b || (a && b && a->entries == b->entries);
}

// ------------------------------------------------------------------ pipeline
const VCol* PipelineBuild::add_col(VCol c) {
  owned.push_back(std::make_unique<VCol>(std::move(c)));
  return owned.back().get();
}

NodePtr PipelineBuild::resolve(const ml_expr& e) const { return resolve_layout(parse_expr(e), layout); }

void PipelineBuild::source(const ml_batch* b, bool use_selection) {
  if (!b) fail(ML_ERR_INTERNAL, "null batch");
  layout.clear();
  for (int i = 0; i < b->num_columns; i++) {
    const ml_column& mc = b->columns[i];
    if (mc.len != b->num_rows) fail(ML_ERR_INTERNAL, "column length differs from batch rows");
    if (mc.type < ML_TYPE_BOOLEAN || mc.type > ML_TYPE_UTF8) fail(ML_ERR_PLAN, "bad column type");
    VCol c;
    c.type = mc.type;
    c.stage = 0;
    c.data = mc.data;
    if (mc.type == ML_TYPE_UTF8) {
      if (!mc.dict) fail(ML_ERR_UNSUPPORTED, "UTF8 columns must be dictionary-encoded");
      c.dict = ctx->dict(mc.dict);
    }
    if (b->num_rows > 0 && !mc.data) fail(ML_ERR_INTERNAL, "null column data");
    if (mc.validity && b->num_rows > 0) {
      // the pipeline reads validity as one byte per physical row
      if (bump) {
        uint8_t* v = (uint8_t*)bump->take((size_t)b->num_rows);
        launch_bits_to_bytes(mc.validity, b->num_rows, v, ctx->stream);
        c.valid = v;
      } else {
        Buffer v(ctx->arena, (size_t)b->num_rows);
        launch_bits_to_bytes(mc.validity, b->num_rows, v.as<uint8_t>(), ctx->stream);
        c.valid = v.as<uint8_t>();
        tables.push_back(std::move(v));
      }
    }
    layout.push_back(add_col(std::move(c)));
  }
  P.sel = (use_selection && b->selection) ? b->selection : nullptr;
  active_rows = P.sel ? b->selection_len : b->num_rows;
  P.begin = 0;
  P.end = active_rows;
  P.expand_step = -1;
  P.nsteps = 0;
}

void PipelineBuild::add_ops(const ml_operator* ops, int n) {
  int ntables = 0;
  for (int i = 0; i < n; i++) {
    const ml_operator& op = ops[i];
    switch (op.kind) {
      case ML_PIPE_FILTER: {
        NodePtr pred = resolve(op.predicate);
        if (comp.type_of(pred) != ML_TYPE_BOOLEAN) fail(ML_ERR_EXECUTION, "filter predicate is not boolean");
        if (P.nsteps >= MAX_STEPS) fail(ML_ERR_UNSUPPORTED, "too many operators in one pipeline");
        Step& st = P.steps[P.nsteps++];
        st = Step{};
        st.kind = STEP_FILTER;
        st.pred = comp.compile(pred, -1);
        break;
      }
      case ML_PIPE_PROBE: {
        const ml_join_table* t = op.table;
        if (!t) fail(ML_ERR_INTERNAL, "probe before build");
        join_table_resolve(const_cast<ml_join_table*>(t));  // the layout is final from here on
        const JoinTableImpl& J = t->impl;
        if (op.num_keys != (int)J.key_types.size()) fail(ML_ERR_PLAN, "probe and build key counts differ");
        if (P.nsteps >= MAX_STEPS || nstages >= MAX_STAGES) fail(ML_ERR_UNSUPPORTED, "too many joins in one pipeline");
        const int stage = nstages++;
        Step& st = P.steps[P.nsteps];
        st = Step{};
        st.kind = STEP_PROBE;
        st.stage = stage;
        st.table = ntables;
        st.nkeys = op.num_keys;
        NodePtr key0;
        for (int k = 0; k < op.num_keys; k++) {
          NodePtr kn = resolve(op.keys[k]);
          if (k == 0) key0 = kn;
          if (J.key_types[k] == ML_TYPE_UTF8) {
            const VCol* c = kn->kind == ML_EXPR_COLUMN ? kn->col : nullptr;
            if (!c || !same_dict(c->dict, J.key_dicts[k]))
              fail(ML_ERR_UNSUPPORTED, "string join keys must share the build side's dictionary");
          }
          st.keys[k] = comp.compile_join_key(kn, J.key_types[k]);
        }
        P.tables[ntables++] = J.dev;
        if (!J.dev.unique) {
          if (expanding) fail(ML_ERR_UNSUPPORTED, "more than one join with duplicate build keys in one pipeline");
          expanding = true;
          P.expand_step = P.nsteps;
        }
        P.nsteps++;
        std::vector<const VCol*> build;
        for (size_t c = 0; c < J.col_types.size(); c++) {
          VCol v;
          v.type = J.col_types[c];
          v.stage = stage;
          v.data = J.cols[c].ptr;
          v.valid = J.col_valid[c].ptr ? J.col_valid[c].as<uint8_t>() : nullptr;
          v.dict = J.col_dicts[c];
          if ((int)c == J.implied_col) {
            // the table's first-key column, not stored: a matched row's value
            // is the probe key (equal key words), in the column's type
            NodePtr e = key0;
            if (comp.type_of(key0) != v.type) {
              auto cast = std::make_shared<Node>();
              cast->kind = ML_EXPR_CAST;
              cast->type = v.type;
              cast->left = key0;
              e = cast;
            }
            v.data = nullptr;
            v.valid = nullptr;
            v.expr = e;
          }
          build.push_back(add_col(std::move(v)));
        }
        // output layout: build ++ probe or probe ++ build (join.rs:331-335)
        std::vector<const VCol*> nl;
        if (op.build_is_left) {
          nl = build;
          nl.insert(nl.end(), layout.begin(), layout.end());
        } else {
          nl = layout;
          nl.insert(nl.end(), build.begin(), build.end());
        }
        layout = nl;
        break;
      }
      case ML_PIPE_PROJECT: {
        std::vector<const VCol*> nl;
        for (int k = 0; k < op.num_exprs; k++) {
          NodePtr e = resolve(op.exprs[k]);
          VCol v;
          v.type = comp.type_of(e);
          if (v.type == ML_TYPE_NULL) v.type = ML_TYPE_INT64;
          v.expr = e;
          if (e->kind == ML_EXPR_COLUMN) {
            v.dict = e->col->dict;
            v.stage = e->col->stage;
          }
          nl.push_back(add_col(std::move(v)));
        }
        layout = nl;
        break;
      }
      default: fail(ML_ERR_PLAN, "unknown operator kind");
    }
  }
}

void PipelineBuild::upload() {
  if (ins.empty()) ins.push_back(Instr{});
  Instr* dst;
  if (bump) {
    dst = (Instr*)bump->take(ins.size() * sizeof(Instr));
    mlg::upload(ctx, dst, ins.data(), ins.size() * sizeof(Instr));
  } else {
    dev_ins = ctx->const_upload(ins.data(), ins.size() * sizeof(Instr));
    dst = dev_ins->as<Instr>();
  }
  P.ins = dst;
  P.nins = (int32_t)ins.size();
}

int PipelineBuild::grid(int block) const {
  int64_t g = (P.end - P.begin + block - 1) / block;
  const int per_cu = ctx->cfg.grid_per_cu > 0 ? ctx->cfg.grid_per_cu : 8;
  const int64_t cap = (int64_t)ctx->num_cus * per_cu;
  if (g > cap) g = cap;
  if (g < 1) g = 1;
  return (int)g;
}

#ifndef ML_BLOOM_KEYS_PER_WORD
#define ML_BLOOM_KEYS_PER_WORD 8
#endif
#ifndef ML_JOIN_LOAD_PCT
#define ML_JOIN_LOAD_PCT 75
#endif

static int block_of(ml_ctx* ctx) { return ctx->cfg.block_size > 0 ? ctx->cfg.block_size : 256; }

// Stages whose row index a program's loads read.
static uint32_t prog_stages(const std::vector<Instr>& ins, const Prog& p) {
  uint32_t m = 0;
  for (int i = 0; i < p.count; i++) {
    const Instr& in = ins[(size_t)p.first + i];
    if (reads_column(in.op)) m |= 1u << in.stage;
  }
  return m;
}

static uint32_t step_stages(const std::vector<Instr>& ins, const Step& st) {
  if (st.kind == STEP_FILTER) return prog_stages(ins, st.pred);
  uint32_t m = 0;
  for (int k = 0; k < st.nkeys; k++) m |= prog_stages(ins, st.keys[k]);
  return m;
}

struct FlowLaunch {
  int grid = 0;
  size_t lds = 0;
  int64_t nwaves = 0;
};

constexpr int FLOW_BLOCK = 256;
constexpr int64_t FLOW_MIN_CHUNK = 512;  // source rows per wave at least
constexpr size_t FLOW_MAX_LDS = 64 * 1024;

// Prepare B.P for flow_kernel: cut the steps into segments (a queue after
// every step, or after every probe when there are more steps than levels),
// keep in each queue only the stages read further down, and size the waves
// so every CU is filled. `sink_stages` are the stages the sink reads. Returns
// false when the general kernel must run instead (a probe that can match
// several rows, row indices beyond 32 bits, or queues too large for LDS).
static bool plan_flow(PipelineBuild& B, int sink, uint32_t sink_stages, size_t cells_bytes, FlowLaunch& F,
                      bool sink_at_level0 = false) {
  Plan& P = B.P;
  P.nseg = 0;
  const int64_t active = P.end - P.begin;
  if (B.expanding || P.expand_step >= 0 || active <= 0) return false;
  if (P.end > (int64_t)UINT32_MAX) return false;
  std::vector<int> cuts;
  for (int s = 1; s <= P.nsteps; s++) cuts.push_back(s);
  // a lone filter in front of a sink that evaluates a lane's rows together:
  // the sink runs in the filter's round (no queue before an empty segment)
  if (sink_at_level0 && P.nsteps == 1 && P.steps[0].kind == STEP_FILTER) cuts.clear();
  if ((int)cuts.size() > MAX_LEVELS) {
    std::vector<int> c2;
    for (int s = 1; s < P.nsteps; s++)
      if (P.steps[s - 1].kind == STEP_PROBE) c2.push_back(s);
    c2.push_back(P.nsteps);
    if ((int)c2.size() > MAX_LEVELS) c2.erase(c2.begin() + (MAX_LEVELS - 1), c2.end() - 1);
    cuts = c2;
  }
  const int nseg = (int)cuts.size() + 1;
  // streaming first probe: step 0 alone in segment 0, probing a unique table
  // with plain non-null stage-0 column loads as key words
  P.fast = 0;
  if (nseg >= 2 && cuts[0] == 1 && P.steps[0].kind == STEP_PROBE && P.steps[0].nkeys <= 2 &&
      P.tables[P.steps[0].table].unique) {
    bool plain = true;
    for (int k = 0; k < P.steps[0].nkeys; k++) {
      const Prog& kp = P.steps[0].keys[k];
      const Instr& in = B.ins[(size_t)kp.first];
      plain = plain && kp.count == 1 && !kp.nullable && in.op < OP_CONST && in.stage == 0 && in.imm2 == 0;
      if (plain) {
        P.fast_op[k] = in.op;
        P.fast_col[k] = (const void*)in.imm;
      }
    }
    if (plain) {
      P.fast = 1;
      P.fast_nkeys = P.steps[0].nkeys;
    }
  }
  // a segment-0 filter made only of column atoms: truth table over the atoms
  P.natoms = 0;
  if (!P.fast && (cuts.empty() ? P.nsteps : cuts[0]) == 1 && P.steps[0].kind == STEP_FILTER &&
      !P.steps[0].pred.nullable) {
    const Prog& pg = P.steps[0].pred;
    std::vector<FilterAtom> atoms;
    bool ok = true;
    for (int i = 0; i < pg.count && ok; i++) {
      const Instr& in = B.ins[(size_t)pg.first + i];
      if (in.op == OP_CMPC || in.op == OP_LOOKUPC) {
        ok = in.stage == 0 && atoms.size() < (size_t)MAX_ATOMS;
        atoms.push_back(FilterAtom{(const void*)in.imm, in.imm2, in.op, in.flags});
      } else {
        ok = in.op == OP_AND || in.op == OP_OR || in.op == OP_NOT || (in.op == OP_CONST && (in.flags & 1u));
      }
    }
    if (ok && !atoms.empty()) {
      uint32_t truth = 0;
      for (uint32_t m = 0; m < (1u << atoms.size()); m++) {
        std::vector<int> st;
        int a = 0;
        for (int i = 0; i < pg.count; i++) {
          const Instr& in = B.ins[(size_t)pg.first + i];
          if (in.op == OP_CMPC || in.op == OP_LOOKUPC) st.push_back((m >> a++) & 1);
          else if (in.op == OP_CONST) st.push_back(in.imm != 0);
          else if (in.op == OP_NOT) st.back() = !st.back();
          else {
            const int y = st.back();
            st.pop_back();
            st.back() = in.op == OP_AND ? (st.back() && y) : (st.back() || y);
          }
        }
        if (st.size() == 1 && st[0]) truth |= 1u << m;
      }
      P.natoms = (int32_t)atoms.size();
      P.atom_truth = truth;
      for (size_t k = 0; k < atoms.size(); k++) P.atoms[k] = atoms[k];
    }
  }
  // otherwise the stage-0 columns segment 0 loads (at most two) are loaded
  // ahead for all rows of a level-0 round
  P.pf_n = 0;
  if (!P.fast) {
    const int end0 = cuts.empty() ? P.nsteps : cuts[0];
    for (int st = 0; st < end0 && P.pf_n < 2; st++) {
      const Step& S = P.steps[st];
      const int np = S.kind == STEP_FILTER ? 1 : S.nkeys;
      for (int k = 0; k < np && P.pf_n < 2; k++) {
        const Prog& pg = S.kind == STEP_FILTER ? S.pred : S.keys[k];
        for (int i = 0; i < pg.count && P.pf_n < 2; i++) {
          const Instr& in = B.ins[(size_t)pg.first + i];
          if (!reads_column(in.op) || in.stage != 0) continue;
          bool seen = false;
          for (int j = 0; j < P.pf_n; j++) seen = seen || P.pf_col[j] == (const void*)in.imm;
          if (!seen) {
            P.pf_col[P.pf_n] = (const void*)in.imm;
            P.pf_op[P.pf_n] = (uint16_t)column_load_op(in.op, in.flags);
            P.pf_n++;
          }
        }
      }
    }
  }
  uint32_t used_from[MAX_STEPS + 1];
  used_from[P.nsteps] = sink_stages;
  for (int s = P.nsteps - 1; s >= 0; s--) used_from[s] = used_from[s + 1] | step_stages(B.ins, P.steps[s]);
  uint32_t assigned = 1;
  int qw = 0;
  int prev = 0;
  for (int l = 0; l < nseg; l++) {
    const int end = l < (int)cuts.size() ? cuts[l] : P.nsteps;
    P.seg_end[l] = end;
    if (l >= 1) {
      const uint32_t live = assigned & used_from[prev];
      P.live[l] = live;
      P.qoff[l] = qw;
      P.qcap[l] = l == 1 ? QCAP_FAST : QCAP;
      qw += __builtin_popcount(live) * P.qcap[l];
    }
    for (int s = prev; s < end; s++)
      if (P.steps[s].kind == STEP_PROBE) assigned |= 1u << P.steps[s].stage;
    prev = end;
  }
  const size_t lds = cells_bytes + (size_t)(FLOW_BLOCK / 64) * qw * 4;
  if (lds > FLOW_MAX_LDS) return false;
  P.nseg = nseg;
  P.qwords = qw;
  // the stages a ROWS sink records
  P.tup_mask = sink_stages;
  // one wave per resident slot, each with a contiguous chunk of source rows
#ifdef ML_FLOW_BPC
  const int bpc = ML_FLOW_BPC;  // a fixed workgroups-per-CU count (occupancy sweeps)
#else
  const int bpc = flow_blocks_per_cu(sink, P.fast != 0, FLOW_BLOCK, lds);
#endif
#ifdef ML_FLOW_DEBUG
  fprintf(stderr, "[flow] sink %d fast %d lds %zu blocks/CU %d cus %d rows %lld\n", sink, (int)P.fast, lds, bpc,
          B.ctx->num_cus, (long long)active);
#endif
  int64_t nwaves = (int64_t)B.ctx->num_cus * bpc * (FLOW_BLOCK / 64);
  int64_t chunk = (active + nwaves - 1) / nwaves;
  if (chunk < FLOW_MIN_CHUNK) {
    nwaves = (active + FLOW_MIN_CHUNK - 1) / FLOW_MIN_CHUNK;
    nwaves = (nwaves + 3) / 4 * 4;
    chunk = (active + nwaves - 1) / nwaves;
  }
  P.chunk = chunk;
  P.staging = (uint64_t)nwaves * (uint64_t)chunk;
  F.grid = (int)(nwaves / 4);
  F.lds = lds;
  F.nwaves = nwaves;
  return true;
}

// Stages read by the sink programs of a pipeline's aggregate / materialization.
static uint32_t sink_read_stages(const PipelineBuild& B) {
  const Plan& P = B.P;
  uint32_t m = 0;
  for (int k = 0; k < P.nkeys; k++) m |= prog_stages(B.ins, P.keys[k]);
  for (int a = 0; a < P.nargs; a++) m |= prog_stages(B.ins, P.args[a]);
  for (int c = 0; c < P.ncols; c++) m |= prog_stages(B.ins, P.cols[c]);
  if (P.pair_out)
    for (int s = 0; s < P.pair_stages; s++) m |= 1u << s;
  return m;
}

// Run a ROWS flow and materialize its rows densely (sink = SINK_COLLECT or
// SINK_JOIN_BUILD, whose outputs are already set in B.P). Returns the row count.
static uint64_t run_flow_rows(PipelineBuild& B, int sink, const FlowLaunch& F, Buffer* keep_offsets = nullptr) {
  ml_ctx* ctx = B.ctx;
  Buffer tuples(ctx->arena, std::max<uint64_t>(1, (uint64_t)__builtin_popcount(B.P.tup_mask) * B.P.staging) * 4);
  Buffer counts(ctx->arena, (size_t)F.nwaves * 4);
  Buffer offsets(ctx->arena, (size_t)(F.nwaves + 1) * 8);
  B.P.tuples = tuples.as<uint32_t>();
  B.P.wave_counts = counts.as<uint32_t>();
  const int keep = B.P.sink;
  B.P.sink = SINK_ROWS;
  ctx->launch(B.P, F.grid, FLOW_BLOCK, F.lds);
  B.P.sink = keep;
  launch_scan_counts(counts.as<uint32_t>(), F.nwaves, offsets.as<uint64_t>(), ctx->stream);
  launch_materialize(B.P, sink, offsets.as<uint64_t>(), F.nwaves, ctx->stream);
  // both buffers are released on return: a later launch of B must not see them
  B.P.tuples = nullptr;
  B.P.wave_counts = nullptr;
  if (keep_offsets) {  // the caller reads the row count later (offsets[nwaves])
    *keep_offsets = std::move(offsets);
    return UINT64_MAX;
  }
  uint64_t total = 0;
  d2h(ctx, &total, offsets.as<uint64_t>() + F.nwaves, 8);
  return total;
}

// Columns of a layout materialized on the device (collect, join payload).
struct Materialized {
  std::vector<int> types;
  std::vector<std::shared_ptr<HostDict>> dicts;
  std::vector<Buffer> cols;
  std::vector<Buffer> valid;
  int64_t rows = 0;
};

// Compile a layout column for materialization.
static Prog column_prog(PipelineBuild& B, const VCol* c, bool* nullable) {
  Prog p;
  if (c->expr) p = B.comp.compile(c->expr, c->type);
  else p = B.comp.compile_load(*c);
  *nullable = p.nullable != 0;
  return p;
}

// Run B with a COLLECT sink (optionally with per-stage row ids); rows come back in `out`.
static uint64_t run_count(PipelineBuild& B) {
  ml_ctx* ctx = B.ctx;
  uint64_t* counters = ctx->counters.as<uint64_t>();
  B.P.sink = SINK_COUNT_PAIRS;
  B.upload();
  FlowLaunch F;
  if (plan_flow(B, SINK_COUNT_PAIRS, 0, 0, F)) {
    Buffer counts(ctx->arena, (size_t)F.nwaves * 4);
    Buffer offsets(ctx->arena, (size_t)(F.nwaves + 1) * 8);
    B.P.wave_counts = counts.as<uint32_t>();
    ctx->launch(B.P, F.grid, FLOW_BLOCK, F.lds);
    B.P.wave_counts = nullptr;  // released on return
    launch_scan_counts(counts.as<uint32_t>(), F.nwaves, offsets.as<uint64_t>(), ctx->stream);
    uint64_t n = 0;
    d2h(ctx, &n, offsets.as<uint64_t>() + F.nwaves, 8);
    return n;
  }
  B.P.out_count = counters;
  zero(ctx, counters, 16);
  if (B.active_rows > 0) ctx->launch(B.P, B.grid(), block_of(ctx), 0);
  uint64_t n = 0;
  d2h(ctx, &n, counters, 8);
  return n;
}

// Returns false when a capacity hint was too small (nothing usable written).
static bool run_collect(PipelineBuild& B, Materialized& out, bool with_cols, Buffer* pair_rows, int pair_stages,
                        int64_t capacity_hint = -1) {
  ml_ctx* ctx = B.ctx;
  uint64_t* counters = ctx->counters.as<uint64_t>();
  // capacity: active rows, unless one probe can match several build rows
  uint64_t cap = (uint64_t)B.active_rows;
  // row-per-lane collect (below): per-thread output counts, their offsets
  const int grid = B.grid(), block = block_of(ctx);
  const int64_t nthreads = (int64_t)grid * block;
  Buffer tcount, toff;
  uint64_t hc[2] = {0, 0};
  bool counted = false;
  auto count_pass = [&] {  // per-thread counts into tcount, the total into hc
    tcount = Buffer(ctx->arena, (size_t)nthreads * 4);
    toff = Buffer(ctx->arena, (size_t)(nthreads + 1) * 8);
    B.P.sink = SINK_COUNT_PAIRS;
    B.P.out_count = counters;
    zero(ctx, counters, 16);
    B.upload();
    B.P.wave_counts = tcount.as<uint32_t>();
    ctx->launch(B.P, grid, block, 0);
    B.P.wave_counts = nullptr;
    d2h(ctx, hc, counters, 16);
    counted = true;
  };
  if (capacity_hint >= 0) {
    cap = (uint64_t)capacity_hint;
  } else if (B.expanding) {
    // the row-per-lane collect's counting pass sizes the output: the probe
    // chain is walked twice (count, collect), not three times
    count_pass();
    cap = hc[0];
  }
  if (with_cols) {
    if ((int)B.layout.size() > MAX_COLS) fail(ML_ERR_UNSUPPORTED, "too many columns to materialize");
    B.P.ncols = (int)B.layout.size();
    for (size_t c = 0; c < B.layout.size(); c++) {
      const VCol* v = B.layout[c];
      bool nullable = false;
      B.P.cols[c] = column_prog(B, v, &nullable);
      const int w = type_width(v->type);
      B.P.col_width[c] = w;
      out.types.push_back(v->type);
      out.dicts.push_back(v->dict);
      out.cols.emplace_back(ctx->arena, std::max<uint64_t>(cap, 1) * w);
      out.valid.emplace_back();
      if (nullable) out.valid.back() = Buffer(ctx->arena, std::max<uint64_t>(cap, 1));
      B.P.col_out[c] = out.cols.back().ptr;
      B.P.col_valid[c] = out.valid.back().ptr ? out.valid.back().as<uint8_t>() : nullptr;
    }
  } else {
    B.P.ncols = 0;
  }
  if (pair_rows) {
    *pair_rows = Buffer(ctx->arena, std::max<uint64_t>(cap, 1) * pair_stages * 8);
    B.P.pair_out = pair_rows->as<int64_t>();
    B.P.pair_stages = pair_stages;
  } else {
    B.P.pair_out = nullptr;
  }
  B.P.sink = SINK_COLLECT;
  B.P.out_count = counters;
  B.P.out_capacity = cap;
  B.P.out_overflow = (uint32_t*)(counters + 1);
  B.upload();
  FlowLaunch F;
  if (!counted && cap >= (uint64_t)B.active_rows && plan_flow(B, SINK_ROWS, sink_read_stages(B), 0, F)) {
    if (B.P.nsteps == 0) {  // no filter or probe: every active row is an output row
      Buffer offsets;
      run_flow_rows(B, SINK_COLLECT, F, &offsets);
      out.rows = B.active_rows;
      return true;
    }
    out.rows = (int64_t)run_flow_rows(B, SINK_COLLECT, F);
    return true;
  }
  // row-per-lane pipeline: a counting pass stores each thread's output rows,
  // a scan turns them into per-thread offsets, and the collect pass writes
  // every row at its thread's own cursor; one shared counter for all matches
  // would serialize an expanding probe whose matches are scattered over lanes
  if (!counted) count_pass();
  B.P.sink = SINK_COLLECT;
  const bool two_pass = !(hc[1] & 2) && hc[0] <= cap;
  if (two_pass) {
    launch_scan_counts(tcount.as<uint32_t>(), nthreads, toff.as<uint64_t>(), ctx->stream);
    B.P.thread_offsets = toff.as<uint64_t>();
    zero(ctx, counters, 16);
    ctx->launch(B.P, grid, block, 0);
    B.P.thread_offsets = nullptr;
    const uint64_t total = hc[0];
    d2h(ctx, hc, counters, 16);
    hc[0] = total;
  } else if (hc[0] <= cap) {  // a thread's count overflowed 32 bits: one shared counter
    zero(ctx, counters, 16);
    ctx->launch(B.P, grid, block, 0);
    d2h(ctx, hc, counters, 16);
  } else {
    hc[1] |= 1;  // more rows than the capacity
  }
  if (hc[1] & 1) {
    if (capacity_hint >= 0) return false;
    fail(ML_ERR_INTERNAL, "collect output overflowed its capacity");
  }
  out.rows = (int64_t)hc[0];
  return true;
}

static ml_result* make_result(ml_ctx* ctx, Materialized& m) {
  auto* r = new ml_result();
  r->ctx = ctx;
  r->device = ctx->device;
  r->num_rows = m.rows;
  for (size_t c = 0; c < m.types.size(); c++) {
    ResultColumn rc;
    rc.type = m.types[c];
    rc.data = std::move(m.cols[c]);
    rc.valid = std::move(m.valid[c]);
    rc.dict = m.dicts[c];
    r->cols.push_back(std::move(rc));
  }
  r->finish_views();
  return r;
}

// ------------------------------------------------------------- aggregates
struct AggDef {
  int func;
  int arg_type;
  NodePtr arg;  // unresolved (parsed) argument, null for COUNT(*)
  const ml_expr* raw;
};

struct OutDef {
  int out_type;
  int finish;       // FIN_*
  int state;        // slot index
  int count;        // slot whose value > 0 means "valid" (-1: always valid)
  int rank = -1;    // FIN_RANK: index of the aggregate's RankMap
};

// MIN/MAX of a dictionary string column (State::MinMaxStr, accumulator.rs:94,
// 352-368): strings compare bytewise, so the accumulators fold the byte-order
// rank of each row's entry (equal strings share a rank) and the output maps a
// rank back to the code of an entry with that string, in the same dictionary.
struct RankMap {
  std::shared_ptr<HostDict> dict;
  Buffer to_rank;                  // u64 per code (OP_LOOKUP_U64 table)
  Buffer to_code;                  // u32 per rank
  std::vector<uint32_t> host_code; // u32 per rank
};

// Slot "cast" of an argument folded as its dictionary entry's rank.
constexpr int CAST_RANK = 1000;

}  // namespace mlg

struct ml_hashagg {
  ml_ctx* ctx = nullptr;
  std::vector<ml_expr> group_raw;
  std::vector<int> key_types;
  std::vector<mlg::AggDef> aggs;       // argument sources of the slots (Final: one per state column)
  std::vector<mlg::AggDef> user_aggs;  // the aggregates as specified (func, arg_type)
  std::vector<int> state_col;          // Final: index in aggs of each aggregate's first state column
  int mode = ML_AGG_MODE_SINGLE;
  int sizing = ML_SIZING_PLAIN;
  int64_t hint = 0;
  // storage of copied expressions
  std::vector<std::vector<ml_expr_node>> node_store;
  std::deque<std::string> str_store;

  // plan, fixed at the first sink
  bool planned = false;
  bool direct = false;
  bool key_nullable = false;
  std::vector<mlg::AccSlot> slots;   // arg = index into aggs (-1 = rows)
  std::vector<int> slot_cast;        // accumulator input type of the slot's argument
  std::vector<mlg::OutDef> outs;
  std::vector<std::shared_ptr<mlg::HostDict>> key_dicts;
  std::vector<mlg::RankMap> ranks;   // MIN/MAX of strings: one per argument dictionary
  std::vector<int> agg_rank;         // per entry of aggs: its RankMap, or -1
  // direct
  int64_t ngroups = 0;
  std::vector<int64_t> dim_size, dim_stride;
  mlg::Buffer totals;
  // hash
  int nkw = 1;
  uint64_t cap = 0;
  uint64_t wrap = 0;   // probe wrap (cap - 1, or a partition's slots - 1 after a partitioned build)
  mlg::Buffer table;
  std::vector<mlg::Buffer> acc;
  mlg::Buffer flags;   // [0] group count (u64), [1] overflow bits (u32)
  mlg::Buffer nan_ids; // u64 counter numbering NaN group keys (each NaN row is its own group)
  uint64_t len_host = 0;
  bool len_known = true;
  int resizes = 0;
  bool failed = false;
  // memory budget and spill (aggregate/hash.rs:189-282)
  int64_t reserved = 0;           // bytes of the group table held against the context's budget
  bool emit_states = false;       // finalize emits partial states (a spilled Partial aggregate's merge)
  bool no_spill = false;          // a spill merge: over budget is an error
  bool spilled = false;
  std::vector<mlg::OutDef> state_outs;  // partial-state emission (Accumulator::state)
  struct SpillColumn {
    int type;
    std::vector<uint8_t> data, valid;  // valid empty: all rows valid
  };
  struct SpillBatch {
    int64_t rows = 0;
    std::vector<SpillColumn> cols;
  };
  std::vector<std::vector<SpillBatch>> spill;  // per partition, the spilled state batches
  std::shared_ptr<mlg::MemPool> pool;  // the budget `reserved` is held against (outlives the context)
  // LocalSink::combine (hash.rs:439-483): the merged states of combined
  // per-thread aggregates, and the lock combine takes
  ml_hashagg* merger = nullptr;
  bool states_out = false;        // a spilled aggregate's merge emits states (it is being combined)
  std::mutex combine_mu;
  ~ml_hashagg() {
    if (pool) pool->shrink(reserved);
    delete merger;
  }
};

namespace mlg {

static uint64_t next_pow2(uint64_t x) {
  uint64_t p = 1;
  while (p < x) p <<= 1;
  return p;
}

// SwissTable::with_capacity sizing (crates/minilake-hashtable/src/swiss.rs:86-96)
static uint64_t swiss_capacity(uint64_t n) { return std::max<uint64_t>(16, next_pow2(n * 8 / 7 + 1)); }

static bool int_like(int t) { return t == ML_TYPE_INT32 || t == ML_TYPE_INT64 || t == ML_TYPE_DATE; }

// Partial-state column types of one aggregate (Accumulator::state_types,
// accumulator.rs:177-186).
static std::vector<int> agg_state_types(int func, int arg_type) {
  switch (func) {
    case ML_AGG_COUNT_STAR: case ML_AGG_COUNT: return {ML_TYPE_INT64};
    case ML_AGG_SUM:
      if (arg_type == ML_TYPE_FLOAT64) return {ML_TYPE_FLOAT64};
      if (int_like(arg_type) || arg_type == ML_TYPE_BOOLEAN) return {ML_TYPE_INT64};
      break;
    case ML_AGG_AVG:
      if (int_like(arg_type) || arg_type == ML_TYPE_FLOAT64) return {ML_TYPE_FLOAT64, ML_TYPE_INT64};
      break;
    case ML_AGG_MIN: case ML_AGG_MAX:
      if (int_like(arg_type)) return {ML_TYPE_INT64};
      if (arg_type == ML_TYPE_FLOAT64) return {ML_TYPE_FLOAT64};
      if (arg_type == ML_TYPE_UTF8) return {ML_TYPE_UTF8};
      break;
    default: fail(ML_ERR_PLAN, "unknown aggregate function");
  }
  fail(ML_ERR_PLAN, std::string("aggregate of ") + type_name(arg_type) + " not supported");
}

// Final-value type of one aggregate (Accumulator::output_type).
static int agg_final_type(int func, int arg_type) {
  if (func == ML_AGG_COUNT || func == ML_AGG_COUNT_STAR) return ML_TYPE_INT64;
  if (func == ML_AGG_AVG) return ML_TYPE_FLOAT64;
  if (func == ML_AGG_SUM) return arg_type == ML_TYPE_FLOAT64 ? ML_TYPE_FLOAT64 : ML_TYPE_INT64;
  return arg_type;
}

// The RankMap of dictionary d (built and uploaded once per aggregate).
static int rank_map(ml_hashagg* A, const std::shared_ptr<HostDict>& d) {
  for (size_t i = 0; i < A->ranks.size(); i++)
    if (A->ranks[i].dict == d) return (int)i;
  const auto& ent = d->entries;
  std::vector<uint32_t> order(ent.size());
  std::iota(order.begin(), order.end(), 0u);
  // std::string compares as unsigned bytes, as Rust's [u8] ordering does
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return ent[a] < ent[b]; });
  std::vector<uint64_t> rank(std::max<size_t>(ent.size(), 1), 0);
  std::vector<uint32_t> code;
  for (size_t r = 0; r < order.size(); r++) {
    if (r == 0 || ent[order[r]] != ent[order[r - 1]]) code.push_back(order[r]);
    rank[order[r]] = code.size() - 1;
  }
  if (code.empty()) code.push_back(0);
  RankMap m;
  m.dict = d;
  m.to_rank = Buffer(A->ctx->arena, rank.size() * 8);
  upload(A->ctx, m.to_rank.ptr, rank.data(), rank.size() * 8);
  m.to_code = Buffer(A->ctx->arena, code.size() * 4);
  upload(A->ctx, m.to_code.ptr, code.data(), code.size() * 4);
  m.host_code = std::move(code);
  A->ranks.push_back(std::move(m));
  return (int)A->ranks.size() - 1;
}

// The dictionary of a string aggregate argument (a dictionary column).
static std::shared_ptr<HostDict> arg_dict(PipelineBuild& B, const ml_expr* raw) {
  NodePtr n = B.resolve(*raw);
  if (n->kind != ML_EXPR_COLUMN || !n->col->dict) fail(ML_ERR_UNSUPPORTED, "min/max of a string needs a dictionary column");
  return n->col->dict;
}

// Final mode (HashAggState::merge_partial, hash.rs:125-137; Accumulator::merge,
// accumulator.rs:407-497): every state column is folded by the accumulator
// that merges it. COUNT adds counts (always valid); SUM and MIN/MAX skip
// NULL states and stay NULL without a valid one; AVG adds sums and counts and
// finishes as sum / count (count 0: NULL).
static void plan_final(ml_hashagg* A, PipelineBuild& B) {
  auto add_slot = [&](int kind, int arg, int cast) {
    for (size_t s = 0; s < A->slots.size(); s++)
      if (A->slots[s].kind == kind && A->slots[s].arg == arg && A->slot_cast[s] == cast) return (int)s;
    A->slots.push_back({kind, arg});
    A->slot_cast.push_back(cast);
    return (int)A->slots.size() - 1;
  };
  for (size_t a = 0; a < A->user_aggs.size(); a++) {
    const AggDef& u = A->user_aggs[a];
    const int sc = A->state_col[a];
    const std::vector<int> types = agg_state_types(u.func, u.arg_type);
    bool nullable = false;
    for (size_t j = 0; j < types.size(); j++) {
      std::vector<Instr> tmp;
      std::vector<Buffer> tt;
      Compiler c(B.ctx, tmp, tt);
      int t = -1;
      Prog p = c.compile(B.resolve(*A->aggs[(size_t)sc + j].raw), -1, &t);
      if (t != types[j])
        fail(ML_ERR_INTERNAL, std::string("aggregate merge got a ") + type_name(t) + " state, wants " + type_name(types[j]));
      if (j == 0) nullable = p.nullable != 0;
    }
    OutDef o{agg_final_type(u.func, u.arg_type), FIN_COPY, 0, -1};
    switch (u.func) {
      case ML_AGG_COUNT_STAR: case ML_AGG_COUNT:
        o.state = add_slot(ACC_SUM_I64, sc, ML_TYPE_INT64);
        A->state_outs.push_back(o);
        break;
      case ML_AGG_SUM:
        o.state = add_slot(types[0] == ML_TYPE_FLOAT64 ? ACC_SUM_F64 : ACC_SUM_I64, sc, types[0]);
        o.count = nullable ? add_slot(ACC_COUNT, sc, types[0]) : 0;
        A->state_outs.push_back(o);
        break;
      case ML_AGG_AVG:
        o.finish = FIN_AVG;
        o.state = add_slot(ACC_SUM_F64, sc, ML_TYPE_FLOAT64);
        o.count = add_slot(ACC_SUM_I64, sc + 1, ML_TYPE_INT64);
        A->state_outs.push_back({ML_TYPE_FLOAT64, FIN_COPY, o.state, -1});
        A->state_outs.push_back({ML_TYPE_INT64, FIN_COPY, o.count, -1});
        break;
      default: {  // MIN / MAX
        const bool mn = u.func == ML_AGG_MIN;
        if (types[0] == ML_TYPE_UTF8) {  // a string state: fold its entry's rank
          o.finish = FIN_RANK;
          o.rank = rank_map(A, arg_dict(B, A->aggs[(size_t)sc].raw));
          A->agg_rank[(size_t)sc] = o.rank;
          o.state = add_slot(mn ? ACC_MIN_I64 : ACC_MAX_I64, sc, CAST_RANK);
          o.count = nullable ? add_slot(ACC_COUNT, sc, CAST_RANK) : 0;
          A->state_outs.push_back(o);
          break;
        }
        const bool f = types[0] == ML_TYPE_FLOAT64;
        if (!f) o.finish = u.arg_type == ML_TYPE_INT64 ? FIN_COPY : FIN_I32;
        o.state = add_slot(f ? (mn ? ACC_MIN_F64 : ACC_MAX_F64) : (mn ? ACC_MIN_I64 : ACC_MAX_I64), sc, types[0]);
        o.count = nullable ? add_slot(ACC_COUNT, sc, types[0]) : 0;
        A->state_outs.push_back({types[0], FIN_COPY, o.state, o.count});
        break;
      }
    }
    A->outs.push_back(o);
  }
  if (A->emit_states) A->outs = A->state_outs;
  if ((int)A->slots.size() > MAX_AGG_SLOTS) fail(ML_ERR_UNSUPPORTED, "too many aggregate states");
}

// Plan accumulator slots (accumulator.rs:104-165 state kinds) for the aggregate
// list. Partial mode keeps Single mode's slots and emits the states instead of
// the finished values (Accumulator::state, accumulator.rs:382-404).
static void plan_agg(ml_hashagg* A, PipelineBuild& B) {
  A->state_outs.clear();
  A->slots.clear();
  A->slot_cast.clear();
  A->outs.clear();
  A->agg_rank.assign(A->aggs.size(), -1);
  A->slots.push_back({ACC_COUNT, -1});  // rows of the group
  A->slot_cast.push_back(-1);
  if (A->mode == ML_AGG_MODE_FINAL) {
    plan_final(A, B);
    return;
  }
  auto add_slot = [&](int kind, int arg, int cast) {
    for (size_t s = 0; s < A->slots.size(); s++)
      if (A->slots[s].kind == kind && A->slots[s].arg == arg && A->slot_cast[s] == cast) return (int)s;
    A->slots.push_back({kind, arg});
    A->slot_cast.push_back(cast);
    return (int)A->slots.size() - 1;
  };
  for (size_t a = 0; a < A->aggs.size(); a++) {
    const AggDef& d = A->aggs[a];
    bool nullable = false;
    if (d.func != ML_AGG_COUNT_STAR) {
      NodePtr n = B.resolve(*d.raw);
      std::vector<Instr> tmp;
      std::vector<Buffer> tt;
      Compiler c(B.ctx, tmp, tt);
      Prog p = c.compile(n, d.arg_type);
      nullable = p.nullable != 0;
    }
    const int ia = (int)a;
    OutDef o{ML_TYPE_INT64, FIN_COPY, 0, -1};
    switch (d.func) {
      case ML_AGG_COUNT_STAR: o.state = 0; break;
      case ML_AGG_COUNT: o.state = nullable ? add_slot(ACC_COUNT, ia, d.arg_type) : 0; break;
      case ML_AGG_SUM:
        if (d.arg_type == ML_TYPE_INT32 || d.arg_type == ML_TYPE_INT64 || d.arg_type == ML_TYPE_BOOLEAN) {
          o.out_type = ML_TYPE_INT64;
          o.state = add_slot(ACC_SUM_I64, ia, ML_TYPE_INT64);
        } else if (d.arg_type == ML_TYPE_FLOAT64) {
          o.out_type = ML_TYPE_FLOAT64;
          o.state = add_slot(ACC_SUM_F64, ia, ML_TYPE_FLOAT64);
        } else {
          fail(ML_ERR_PLAN, std::string("sum(") + type_name(d.arg_type) + ") not supported");
        }
        o.count = nullable ? add_slot(ACC_COUNT, ia, ML_TYPE_INT64) : 0;
        break;
      case ML_AGG_AVG:
        if (d.arg_type != ML_TYPE_INT32 && d.arg_type != ML_TYPE_INT64 && d.arg_type != ML_TYPE_FLOAT64)
          fail(ML_ERR_PLAN, std::string("avg(") + type_name(d.arg_type) + ") not supported");
        o.out_type = ML_TYPE_FLOAT64;
        o.finish = FIN_AVG;
        o.state = add_slot(ACC_SUM_F64, ia, ML_TYPE_FLOAT64);
        o.count = nullable ? add_slot(ACC_COUNT, ia, ML_TYPE_FLOAT64) : 0;
        break;
      case ML_AGG_MIN: case ML_AGG_MAX: {
        const bool mn = d.func == ML_AGG_MIN;
        if (d.arg_type == ML_TYPE_INT32 || d.arg_type == ML_TYPE_INT64 || d.arg_type == ML_TYPE_DATE) {
          o.out_type = d.arg_type;
          o.finish = d.arg_type == ML_TYPE_INT64 ? FIN_COPY : FIN_I32;
          o.state = add_slot(mn ? ACC_MIN_I64 : ACC_MAX_I64, ia, ML_TYPE_INT64);
        } else if (d.arg_type == ML_TYPE_FLOAT64) {
          o.out_type = ML_TYPE_FLOAT64;
          o.state = add_slot(mn ? ACC_MIN_F64 : ACC_MAX_F64, ia, ML_TYPE_FLOAT64);
        } else if (d.arg_type == ML_TYPE_UTF8) {
          o.out_type = ML_TYPE_UTF8;
          o.finish = FIN_RANK;
          o.rank = rank_map(A, arg_dict(B, d.raw));
          A->agg_rank[(size_t)ia] = o.rank;
          o.state = add_slot(mn ? ACC_MIN_I64 : ACC_MAX_I64, ia, CAST_RANK);
        } else {
          fail(ML_ERR_PLAN, std::string("min/max(") + type_name(d.arg_type) + ") not supported");
        }
        o.count = nullable ? add_slot(ACC_COUNT, ia, o.finish == FIN_RANK ? CAST_RANK : ML_TYPE_INT64) : 0;
        break;
      }
      default: fail(ML_ERR_PLAN, "unknown aggregate function");
    }
    // the partial states of this aggregate (accumulator.rs:382-404)
    if (d.func == ML_AGG_AVG) {  // (sum, count), both always valid
      A->state_outs.push_back({ML_TYPE_FLOAT64, FIN_COPY, o.state, -1});
      A->state_outs.push_back({ML_TYPE_INT64, FIN_COPY, o.count, -1});
    } else if (d.func == ML_AGG_MIN || d.func == ML_AGG_MAX) {  // integer states are Int64, strings stay strings
      OutDef st = o;
      if (st.finish != FIN_RANK) {
        st.finish = FIN_COPY;
        if (st.out_type != ML_TYPE_FLOAT64) st.out_type = ML_TYPE_INT64;
      }
      A->state_outs.push_back(st);
    } else {
      A->state_outs.push_back(o);
    }
    if (A->mode != ML_AGG_MODE_PARTIAL) A->outs.push_back(o);
  }
  if (A->mode == ML_AGG_MODE_PARTIAL) A->outs = A->state_outs;
  if ((int)A->slots.size() > MAX_AGG_SLOTS) fail(ML_ERR_UNSUPPORTED, "too many aggregate states");
}

// Compile keys/args of one sink call into B.P (shared by hash and direct modes).
static void compile_agg(ml_hashagg* A, PipelineBuild& B, bool first) {
  if ((int)A->key_types.size() > MAX_KEYS) fail(ML_ERR_UNSUPPORTED, "more than 4 group keys");
  B.P.nkeys = (int)A->key_types.size();
  bool key_nullable = false;
  std::vector<std::shared_ptr<HostDict>> dicts;
  for (size_t k = 0; k < A->key_types.size(); k++) {
    NodePtr n = B.resolve(A->group_raw[k]);
    std::shared_ptr<HostDict> d;
    if (A->key_types[k] == ML_TYPE_UTF8) {
      if (n->kind != ML_EXPR_COLUMN || !n->col->dict) fail(ML_ERR_UNSUPPORTED, "string group keys must be dictionary columns");
      d = n->col->dict;
    }
    dicts.push_back(d);
    B.P.keys[k] = B.comp.compile_key(n, A->key_types[k], A->nan_ids.as<uint64_t>());
    key_nullable = key_nullable || B.P.keys[k].nullable;
  }
  if (first) {
    A->key_dicts = dicts;
    A->key_nullable = key_nullable;
  } else {
    for (size_t k = 0; k < dicts.size(); k++)
      if (!same_dict(dicts[k], A->key_dicts[k])) fail(ML_ERR_UNSUPPORTED, "every batch must share the group key dictionaries");
    if (key_nullable != A->key_nullable) fail(ML_ERR_INTERNAL, "key nullability changed between batches");
  }
  B.P.key_nullable = A->key_nullable ? 1 : 0;
  // argument programs, one per (aggregate, input cast), deduplicated by code
  std::vector<std::pair<int, int>> arg_keys;  // (agg, cast)
  std::vector<int> arg_prog;                  // index into B.P.args
  std::vector<std::pair<int, int>> prog_span;
  B.P.nargs = 0;
  B.P.nacc = (int)A->slots.size();
  for (size_t s = 0; s < A->slots.size(); s++) {
    const AccSlot& sl = A->slots[s];
    B.P.acc[s].kind = sl.kind;
    if (sl.arg < 0) {
      B.P.acc[s].arg = -1;
      continue;
    }
    const std::pair<int, int> key{sl.arg, A->slot_cast[s]};
    int found = -1;
    for (size_t q = 0; q < arg_keys.size(); q++)
      if (arg_keys[q] == key) found = arg_prog[q];
    if (found < 0) {
      const AggDef& d = A->aggs[(size_t)sl.arg];
      NodePtr x = B.resolve(*d.raw);
      int want = d.arg_type;
      if (A->slot_cast[s] >= 0 && A->slot_cast[s] != CAST_RANK && A->slot_cast[s] != d.arg_type) {
        // evaluate_to_column(arg, arg_type), then prepare_input's cast (accumulator.rs:229-236)
        auto c = std::make_shared<Node>();
        c->kind = ML_EXPR_CAST;
        c->type = d.arg_type;
        c->left = x;
        x = c;
        want = A->slot_cast[s];
      }
      Prog p;
      if (A->slot_cast[s] == CAST_RANK) {  // the byte-order rank of the row's dictionary entry
        const int rk = A->agg_rank[(size_t)sl.arg];
        if (rk < 0) fail(ML_ERR_INTERNAL, "string aggregate without a rank map");
        if (!same_dict(arg_dict(B, d.raw), A->ranks[(size_t)rk].dict))
          fail(ML_ERR_UNSUPPORTED, "every batch must share the dictionary of a string MIN/MAX argument");
        p = B.comp.compile(B.resolve(*d.raw), -1);
        Instr in{};
        in.op = OP_LOOKUP_U64;
        in.depth = 1;
        in.imm = (uint64_t)(uintptr_t)A->ranks[(size_t)rk].to_rank.as<uint64_t>();
        B.ins.push_back(in);
        p.count++;
      } else {
        p = B.comp.compile(x, want);
      }
      // dedupe identical code
      for (int q = 0; q < B.P.nargs && found < 0; q++) {
        const Prog& o = B.P.args[q];
        if (o.count == p.count && o.nullable == p.nullable &&
            std::memcmp(&B.ins[(size_t)o.first], &B.ins[(size_t)p.first], sizeof(Instr) * (size_t)p.count) == 0) {
          found = q;
          B.ins.resize((size_t)p.first);
        }
      }
      if (found < 0) {
        if (B.P.nargs >= MAX_AGG_SLOTS) fail(ML_ERR_UNSUPPORTED, "too many aggregate arguments");
        B.P.args[B.P.nargs] = p;
        found = B.P.nargs++;
      }
      arg_keys.push_back(key);
      arg_prog.push_back(found);
    }
    B.P.acc[s].arg = found;
  }
  // an argument whose code begins with all of the previous argument's code
  // (price * (1 - disc) * (1 + tax) after price * (1 - disc)) is evaluated
  // from that argument's values by the 4-row sinks
  B.P.arg_chain = 0;
  for (int a = 1; a < B.P.nargs; a++) {
    const Prog& pr = B.P.args[a - 1];
    const Prog& cu = B.P.args[a];
    if (cu.count > pr.count && pr.count > 0 && (cu.nullable || !pr.nullable) &&
        std::memcmp(&B.ins[(size_t)cu.first], &B.ins[(size_t)pr.first], sizeof(Instr) * (size_t)pr.count) == 0)
      B.P.arg_chain |= 1u << a;
  }
  // stack slots of the key and argument programs (push instructions write
  // slot depth, the others consume from the top)
  auto slots_of = [&](const Prog& pg) {
    int m = 0;
    for (int i = 0; i < pg.count; i++) {
      const Instr& in = B.ins[(size_t)pg.first + i];
      m = std::max(m, in.op <= OP_LOOKUPC ? in.depth + 1 : (int)in.depth);
    }
    return m;
  };
  int need = 0;
  for (int k = 0; k < B.P.nkeys; k++) need = std::max(need, slots_of(B.P.keys[k]));
  for (int a = 0; a < B.P.nargs; a++) need = std::max(need, slots_of(B.P.args[a]));
  B.P.src_batch = need <= 4 ? 1 : 0;
  // the stage-0 columns those programs read, for a batched level-0 sink: at
  // most 8, each marked in the instructions that read it
  B.P.sp_n = 0;
  if (B.P.src_batch) {
    std::vector<const void*> cols;
    std::vector<int> ops;
    std::vector<size_t> marked;
    bool fits = true;
    auto mark = [&](const Prog& pg) {
      for (int i = 0; i < pg.count && fits; i++) {
        Instr& in = B.ins[(size_t)pg.first + i];
        if (!reads_column(in.op) || in.stage != 0) continue;
        const void* c = (const void*)in.imm;
        const int lop = (int)column_load_op(in.op, in.flags);
        size_t j = 0;
        while (j < cols.size() && !(cols[j] == c && ops[j] == lop)) j++;
        if (j == cols.size()) {
          if (cols.size() == 8) {
            fits = false;
            break;
          }
          cols.push_back(c);
          ops.push_back(lop);
        }
        in.flags = (in.flags & 0x00FFFFFFu) | ((uint32_t)(j + 1) << 24);
        marked.push_back((size_t)pg.first + i);
      }
    };
    for (int k = 0; k < B.P.nkeys; k++) mark(B.P.keys[k]);
    for (int a = 0; a < B.P.nargs; a++) mark(B.P.args[a]);
    if (fits) {
      B.P.sp_n = (int32_t)cols.size();
      for (size_t j = 0; j < cols.size(); j++) {
        B.P.sp_col[j] = cols[j];
        B.P.sp_op[j] = ops[j];
      }
    } else {
      for (size_t m : marked) B.ins[m].flags &= 0x00FFFFFFu;
    }
  }
  B.P.count_slots = 0;
  for (int a = 0; a < MAX_AGG_SLOTS; a++) B.P.arg_slots[a] = 0;
  for (size_t s = 0; s < A->slots.size(); s++) {
    if (B.P.acc[s].arg < 0) B.P.count_slots |= 1u << s;
    else B.P.arg_slots[B.P.acc[s].arg] |= 1u << s;
  }
}

static AccSlotList slot_kinds(const ml_hashagg* A) {
  AccSlotList l{};
  for (size_t s = 0; s < A->slots.size(); s++) l.kind[s] = A->slots[s].kind;
  return l;
}

static void init_acc(ml_hashagg* A, std::vector<Buffer>& acc, uint64_t cap) {
  acc.clear();
  for (size_t s = 0; s < A->slots.size(); s++) {
    acc.emplace_back(A->ctx->arena, cap * 8);
    launch_fill_u64(acc.back().as<uint64_t>(), acc_init_bits(A->slots[s].kind), (int64_t)cap, A->ctx->stream);
  }
}

// Grow the group table to hold `needed` groups (SwissTable::reserve,
// swiss.rs:219-232). The new table is reserved against the context's budget
// first (old and new coexist during the rehash); false = refused.
static bool agg_reserve(ml_hashagg* A, uint64_t needed) {
  ml_ctx* ctx = A->ctx;
  if (A->cap && needed * 8 <= A->cap * 7) return true;
  const uint64_t len = A->len_host;
  const uint64_t ncap = swiss_capacity(std::max(needed, len * 2));
  const int64_t bytes = (int64_t)(ncap * (1 + A->nkw) * 8 + ncap * A->slots.size() * 8);
  if (!A->pool->try_grow(bytes)) return false;
  Buffer nt(ctx->arena, ncap * (1 + A->nkw) * 8);
  zero(ctx, nt.ptr, nt.bytes);
  std::vector<Buffer> nacc;
  init_acc(A, nacc, ncap);
  if (A->cap) {
    AccMove mv{};
    mv.n = (int)A->slots.size();
    for (int s = 0; s < mv.n; s++) {
      mv.src[s] = A->acc[s].ptr;
      mv.dst[s] = nacc[s].ptr;
    }
    launch_agg_rehash(A->table.as<uint64_t>(), A->cap, A->nkw, nt.as<uint64_t>(), ncap - 1, mv, ctx->stream);
    A->resizes++;
  }
  A->table = std::move(nt);
  A->acc = std::move(nacc);
  A->cap = ncap;
  A->wrap = ncap - 1;
  A->pool->shrink(A->reserved);
  A->reserved = bytes;
  return true;
}

static uint64_t agg_len(ml_hashagg* A) {
  if (!A->len_known) {
    uint64_t n = 0;
    d2h(A->ctx, &n, A->flags.ptr, 8);
    A->len_host = n;
    A->len_known = true;
  }
  return A->len_host;
}

static void agg_reset_table(ml_hashagg* A) {
  A->table.reset();
  A->acc.clear();
  A->pool->shrink(A->reserved);
  A->reserved = 0;
  A->cap = 0;
  A->len_host = 0;
  A->len_known = true;
  zero(A->ctx, A->flags.ptr, 32);
}

// Partitioned aggregation of a first batch whose group count is known
// (a capacity hint or the learned estimate): the table is cut into
// partitions of S slots; the surviving rows' key and argument values are
// materialized, scattered into partition order, and each partition is
// aggregated in one workgroup's LDS (a partition with many rows is split into
// chunks whose partial tables are merged), then written out in the table's
// slot layout. No global atomics per row, and no grow-and-rehash. Returns
// false when the pipeline or geometry does not fit (the caller then uses the
// global table).
constexpr uint64_t PART_AGG_MIN_GROUPS = 1 << 16;
constexpr uint64_t PART_AGG_MAX_PARTS = 8192;
constexpr uint32_t PART_CHUNK_ROWS = 1 << 16;

static bool agg_partitioned(ml_hashagg* A, PipelineBuild& B) {
  ml_ctx* ctx = A->ctx;
  Plan& P = B.P;
  const int nkw = A->nkw, nacc = (int)A->slots.size(), nargs = P.nargs, nkeys = P.nkeys;
  if (nkeys + nargs > MAX_COLS || nkeys < 1) return false;
  uint64_t cap = std::max<uint64_t>(1024, next_pow2((uint64_t)A->hint * 2));
  uint64_t S = 2048;
  while (S > 256 && S * (1 + nkw + nacc) * 8 > 65536) S >>= 1;
  S = std::min(S, cap);
  // partitions of S slots each, at most PART_AGG_MAX_PARTS of them
  auto fits = [&](uint64_t c) { return S * (1 + nkw + nacc) * 8 <= 65536 && c / S <= PART_AGG_MAX_PARTS && c < (1ull << 32); };
  if (!fits(cap)) return false;

  // 1. key and argument values of the surviving rows, in source order
  P.ncols = nkeys + nargs;
  std::vector<Buffer> vals, oks;
  for (int c = 0; c < P.ncols; c++) {
    const Prog& pg = c < nkeys ? P.keys[c] : P.args[c - nkeys];
    P.cols[c] = pg;
    P.col_width[c] = 8;
  }
  P.pair_out = nullptr;
  FlowLaunch F;
  if (!plan_flow(B, SINK_ROWS, sink_read_stages(B), 0, F)) {
    P.ncols = 0;
    return false;
  }
  const uint64_t cap_rows = (uint64_t)std::max<int64_t>(1, B.active_rows);
  for (int c = 0; c < P.ncols; c++) {
    vals.emplace_back(ctx->arena, cap_rows * 8);
    oks.emplace_back();
    if (P.cols[c].nullable) oks.back() = Buffer(ctx->arena, cap_rows);
    P.col_out[c] = vals.back().ptr;
    P.col_valid[c] = oks.back().ptr ? oks.back().as<uint8_t>() : nullptr;
  }
  const int64_t n = (int64_t)run_flow_rows(B, SINK_COLLECT, F);
  P.ncols = 0;

  // Steps 2-3 run again with a larger capacity while a learned estimate
  // proves too small (an LDS table filled up: flags u32 2, bit 1); the
  // collected rows are partitioned anew, not collected again. A hinted
  // capacity is never grown.
  const bool grow = A->sizing == ML_SIZING_LEARNED;
  for (;;) {
    const uint64_t nparts = cap / S;
    // 2. partition order
    AggRecords R{};
    R.nkeys = nkeys;
    R.nkw = nkw;
    R.nargs = nargs;
    R.key_nullable = A->key_nullable ? 1 : 0;
    bool args_nullable = false;
    for (int k = 0; k < nkeys; k++) {
      R.key_val[k] = vals[k].as<uint64_t>();
      R.key_ok[k] = oks[k].ptr ? oks[k].as<uint8_t>() : nullptr;
    }
    for (int a = 0; a < nargs; a++) {
      R.arg_val[a] = vals[nkeys + a].as<uint64_t>();
      R.arg_ok[a] = oks[nkeys + a].ptr ? oks[nkeys + a].as<uint8_t>() : nullptr;
      args_nullable = args_nullable || R.arg_ok[a];
    }
    const int part_shift = __builtin_ctzll(S);
    const int nblocks = (int)std::min<int64_t>(256, std::max<int64_t>(1, (n + 16383) / 16384));
    const int64_t rpb = std::max<int64_t>(1, (n + nblocks - 1) / nblocks);
    const int64_t nh = (int64_t)nparts * nblocks;
    Buffer hist(ctx->arena, (size_t)nh * 4), offsets(ctx->arena, (size_t)(nh + 1) * 4);
    Buffer sums(ctx->arena, (size_t)(scan_u32_blocks(nh) + 1) * 4);
    Buffer rkeys(ctx->arena, (size_t)std::max<int64_t>(n, 1) * nkw * 8);
    Buffer rargs(ctx->arena, (size_t)std::max<int64_t>(n, 1) * std::max(nargs, 1) * 8);
    Buffer rvalid;
    if (args_nullable) rvalid = Buffer(ctx->arena, (size_t)std::max<int64_t>(n, 1) * 4);
    launch_agg_part_hist(R, n, cap - 1, part_shift, nblocks, rpb, (int)nparts, hist.as<uint32_t>(), ctx->stream);
    launch_scan_u32(hist.as<uint32_t>(), nh, offsets.as<uint32_t>(), sums.as<uint32_t>(), ctx->stream);
    launch_agg_part_scatter(R, n, cap - 1, part_shift, nblocks, rpb, (int)nparts, offsets.as<uint32_t>(),
                            rkeys.as<uint64_t>(), rargs.as<uint64_t>(), rvalid.as<uint32_t>(), ctx->stream);
    // partition starts (every nblocks-th offset, and the total)
    Buffer starts(ctx->arena, (size_t)(nparts + 1) * 4);
    launch_strided_u32(offsets.as<uint32_t>(), nblocks, (int64_t)nparts + 1, starts.as<uint32_t>(), ctx->stream);
    std::vector<uint32_t> st((size_t)nparts + 1);
    d2h(ctx, st.data(), starts.ptr, st.size() * 4);

    // 3. chunks: one per partition, several for partitions with many rows
    std::vector<PartChunk> chunks;
    std::vector<uint32_t> heavy, first;
    uint32_t q = 0;
    for (uint32_t p = 0; p < nparts; p++) {
      const uint32_t lo = st[p], hi = st[p + 1];
      if (hi - lo <= PART_CHUNK_ROWS) {
        chunks.push_back(PartChunk{p, lo, hi, 0xFFFFFFFFu});
        continue;
      }
      heavy.push_back(p);
      first.push_back(q);
      for (uint32_t c = lo; c < hi; c += PART_CHUNK_ROWS) chunks.push_back(PartChunk{p, c, std::min(hi, c + PART_CHUNK_ROWS), q++});
    }
    first.push_back(q);
    A->table = Buffer(ctx->arena, cap * (1 + nkw) * 8);
    init_acc(A, A->acc, cap);
    PartAggPlan G{};
    G.S = (int32_t)S;
    G.nkw = nkw;
    G.nargs = nargs;
    G.nacc = nacc;
    G.mask = cap - 1;
    for (int s2 = 0; s2 < nacc; s2++) {
      G.acc[s2] = P.acc[s2];  // the compiled slots: `arg` indexes the (deduplicated) argument programs
      G.acc_data[s2] = A->acc[s2].ptr;
    }
    G.slots = A->table.as<uint64_t>();
    G.flags = (uint32_t*)A->flags.ptr;
    Buffer dchunks(ctx->arena, chunks.size() * sizeof(PartChunk));
    h2d(ctx, dchunks.ptr, chunks.data(), chunks.size() * sizeof(PartChunk));
    Buffer partial(ctx->arena, std::max<size_t>(1, (size_t)q * S * (1 + nkw + nacc) * 8));
    Buffer dheavy(ctx->arena, std::max<size_t>(1, heavy.size()) * 4), dfirst(ctx->arena, first.size() * 4);
    if (!heavy.empty()) {
      h2d(ctx, dheavy.ptr, heavy.data(), heavy.size() * 4);
      h2d(ctx, dfirst.ptr, first.data(), first.size() * 4);
    }
    launch_agg_part_chunks(G, dchunks.as<PartChunk>(), (int)chunks.size(), rkeys.as<uint64_t>(), rargs.as<uint64_t>(),
                           rvalid.as<uint32_t>(), n, partial.as<uint64_t>(), ctx->stream);
    launch_agg_part_merge(G, dheavy.as<uint32_t>(), dfirst.as<uint32_t>(), (int)heavy.size(), partial.as<uint64_t>(),
                          ctx->stream);
    A->cap = cap;
    A->wrap = S - 1;
    A->len_known = false;
    if (!grow || !fits(cap * 2)) break;
    // groups written; overflow bits | filled partitions << 32; their projected groups
    uint64_t f[3] = {0, 0, 0};
    d2h(ctx, f, A->flags.ptr, 24);
    if (!(f[1] & 2u)) break;
    // An LDS table filled up: everything from the partitioning on again, from
    // an empty table sized for the groups the attempt saw (the partitions that
    // held, exactly; a filled one projected from how far into its records it
    // got, an overestimate when new keys thin out as keys repeat) with the
    // initial 2x margin, and at least twice the capacity; at most the largest
    // partitioned table, and the plain path once that one has filled up.
    const uint64_t over = std::min<uint64_t>(f[1] >> 32, nparts);
    const uint64_t seen = f[0] - std::min<uint64_t>(f[0], over * part_fill_limit((uint32_t)S)) + f[2];
    uint64_t next = std::max<uint64_t>(cap * 2, next_pow2(2 * seen + 1));
    while (next > cap * 2 && !fits(next)) next >>= 1;
    zero(ctx, A->flags.ptr, 32);
    cap = next;
    A->resizes++;
  }
  hip_check(hipStreamSynchronize(ctx->stream), "partitioned aggregation");
  return true;
}

// One sink call: run the pipeline in B into the aggregate.
static ml_result* agg_finalize(ml_hashagg* A, bool states);

// MiniLakeError::ResourcesExhausted (memory.rs:110-121)
[[noreturn]] static void agg_exhausted(ml_hashagg* A, uint64_t rows) {
  const ml_ctx* ctx = A->ctx;
  fail(ML_ERR_RESOURCES_EXHAUSTED, "HashAggregate: the group table for " + std::to_string(rows) +
                                       " more rows does not fit the memory budget (" + std::to_string(ctx->mem->used) +
                                       " of " + std::to_string(ctx->mem->limit) + " bytes reserved)");
}

constexpr int SPILL_PARTITIONS = 16;  // hash.rs:189-199

// Spill (Inner::spill, hash.rs:220-277): the table's groups as partial states,
// cut into 16 partitions by bits 32..35 of hash_columns over the group keys,
// copied to host memory; the table is emptied and its reservation released.
static void agg_spill(ml_hashagg* A) {
  ml_ctx* ctx = A->ctx;
  for (const OutDef& o : A->state_outs)
    if (o.finish == FIN_RANK) fail(ML_ERR_UNSUPPORTED, "spilling the states of MIN/MAX of strings");
  if (A->cap && agg_len(A) > 0) {
    std::unique_ptr<ml_result> r(agg_finalize(A, true));
    const int64_t n = r->num_rows;
    const int nk = (int)A->key_types.size();
    std::vector<ml_expr_node> nodes((size_t)nk);
    std::vector<ml_expr> keys((size_t)nk);
    for (int k = 0; k < nk; k++) {
      nodes[(size_t)k] = ml_expr_node{};
      nodes[(size_t)k].kind = ML_EXPR_COLUMN;
      nodes[(size_t)k].column = k;
      keys[(size_t)k].nodes = &nodes[(size_t)k];
      keys[(size_t)k].num_nodes = 1;
    }
    ml_batch b{};
    b.columns = r->views.data();
    b.num_columns = nk;
    b.num_rows = n;
    Buffer hd(ctx->arena, (size_t)std::max<int64_t>(n, 1) * 8);
    const ml_status st = ml_hash_columns(ctx, keys.data(), nk, &b, hd.as<uint64_t>());
    if (st != ML_OK) fail(st, ctx->last_error);
    std::vector<uint64_t> h((size_t)n);
    d2h(ctx, h.data(), hd.ptr, (size_t)n * 8);
    std::vector<ml_hashagg::SpillColumn> host;
    for (auto& c : r->cols) {
      ml_hashagg::SpillColumn hc;
      hc.type = c.type;
      hc.data.resize((size_t)n * type_width(c.type));
      d2h(ctx, hc.data.data(), c.data.ptr, hc.data.size());
      if (c.valid.ptr) {
        hc.valid.resize((size_t)n);
        d2h(ctx, hc.valid.data(), c.valid.ptr, (size_t)n);
      }
      host.push_back(std::move(hc));
    }
    if (A->spill.empty()) A->spill.resize(SPILL_PARTITIONS);
    std::vector<std::vector<int64_t>> rows(SPILL_PARTITIONS);
    for (int64_t i = 0; i < n; i++) rows[(size_t)((h[(size_t)i] >> 32) & (SPILL_PARTITIONS - 1))].push_back(i);
    for (int p = 0; p < SPILL_PARTITIONS; p++) {
      if (rows[(size_t)p].empty()) continue;
      ml_hashagg::SpillBatch sb;
      sb.rows = (int64_t)rows[(size_t)p].size();
      for (auto& hc : host) {
        ml_hashagg::SpillColumn pc;
        pc.type = hc.type;
        const size_t w = (size_t)type_width(hc.type);
        pc.data.resize((size_t)sb.rows * w);
        if (!hc.valid.empty()) pc.valid.resize((size_t)sb.rows);
        for (size_t j = 0; j < rows[(size_t)p].size(); j++) {
          const size_t i = (size_t)rows[(size_t)p][j];
          std::memcpy(&pc.data[j * w], &hc.data[i * w], w);
          if (!hc.valid.empty()) pc.valid[j] = hc.valid[i];
        }
        sb.cols.push_back(std::move(pc));
      }
      A->spill[(size_t)p].push_back(std::move(sb));
    }
  }
  agg_reset_table(A);
  A->spilled = true;
}

static void agg_run(ml_hashagg* A, PipelineBuild& B) {
  ml_ctx* ctx = A->ctx;
  const bool first = !A->planned;
  if (first) {
    plan_agg(A, B);
    // direct (small-domain) aggregation when every key is a dictionary or
    // boolean column, and without keys (one group: UngroupedAggregate)
    bool direct = true;
    int64_t groups = 1;
    std::vector<int64_t> sizes;
    for (size_t k = 0; k < A->key_types.size() && direct; k++) {
      NodePtr n = B.resolve(A->group_raw[k]);
      std::vector<Instr> tmp;
      std::vector<Buffer> tt;
      Compiler c(ctx, tmp, tt);
      const bool nl = c.compile_key(n, A->key_types[k]).nullable != 0;
      int64_t dom;
      if (A->key_types[k] == ML_TYPE_UTF8 && n->kind == ML_EXPR_COLUMN && n->col->dict) dom = (int64_t)n->col->dict->entries.size();
      else if (A->key_types[k] == ML_TYPE_BOOLEAN) dom = 2;
      else { direct = false; break; }
      dom += nl ? 1 : 0;
      sizes.push_back(std::max<int64_t>(dom, 1));
      groups *= sizes.back();
      if (groups > 4096) direct = false;
    }
    if (direct && groups * (int64_t)A->slots.size() * 8 > 48 * 1024) direct = false;
    A->direct = direct;
    if (direct) {
      A->ngroups = groups;
      A->dim_size = sizes;
      A->dim_stride.assign(sizes.size(), 1);
      for (int k = (int)sizes.size() - 2; k >= 0; k--) A->dim_stride[k] = A->dim_stride[k + 1] * sizes[k + 1];
      A->totals = Buffer(ctx->arena, (size_t)groups * A->slots.size() * 8);
      AccSlotList kinds = slot_kinds(A);
      std::vector<uint64_t> init((size_t)groups * A->slots.size());
      for (size_t c = 0; c < init.size(); c++) init[c] = acc_init_bits(kinds.kind[c % A->slots.size()]);
      upload(ctx, A->totals.ptr, init.data(), init.size() * 8);
    }
    A->flags = Buffer(ctx->arena, 32);
    zero(ctx, A->flags.ptr, 32);
    A->nan_ids = Buffer(ctx->arena, 8);  // numbers of NaN group keys, never reused
    zero(ctx, A->nan_ids.ptr, 8);
    A->planned = true;
  }
  compile_agg(A, B, first);
  if (first) A->nkw = (int)A->key_types.size() + (A->key_nullable ? 1 : 0);

  if (A->direct) {
    B.P.sink = SINK_DIRECT_AGG;
    for (size_t k = 0; k < A->dim_size.size(); k++) {
      B.P.dim_size[k] = A->dim_size[k];
      B.P.dim_stride[k] = A->dim_stride[k];
    }
    B.P.ngroups = (int32_t)A->ngroups;
    const int ncells = (int)(A->ngroups * (int64_t)A->slots.size());
    B.upload();
    FlowLaunch F;
    int grid, block;
    size_t lds;
    // copies of the cells for the flow kernel (each lane adds to copy lane mod
    // copies): as many as fit 32 KB, or 8 KB when the pipeline probes (its
    // queues and registers bound occupancy, and few rows reach the sink)
    bool probes = false;
    for (int st = 0; st < B.P.nsteps; st++) probes = probes || B.P.steps[st].kind == STEP_PROBE;
    const int cstride = ncells | 1;
    const int budget = probes ? 8192 : 32768;
    int copies = 64;
    while (copies > 1 && (size_t)copies * cstride * 8 > (size_t)budget) copies >>= 1;
    bool flow = false;
    for (; copies >= 1 && !flow; copies >>= 1) {
      B.P.direct_copies = copies;
      B.P.direct_stride = cstride;
      flow = plan_flow(B, SINK_DIRECT_AGG, sink_read_stages(B), (size_t)copies * cstride * 8, F,
                       B.P.src_batch != 0);
      if (!flow && B.P.nseg == 0 && (B.expanding || B.P.expand_step >= 0)) break;
    }
    if (flow) {
      grid = F.grid;
      block = FLOW_BLOCK;
      lds = F.lds;
    } else {
      grid = B.grid();
      block = block_of(ctx);
      lds = (size_t)ncells * 8;
    }
    Buffer partials(ctx->arena, (size_t)grid * ncells * 8);
    B.P.direct_partials = partials.ptr;
    ctx->launch(B.P, grid, block, lds);
    launch_direct_reduce(partials.as<uint64_t>(), grid, ncells, (int)A->slots.size(), slot_kinds(A),
                         A->totals.as<uint64_t>(), ctx->stream);
    return;
  }

  // hash mode
  B.P.sink = SINK_HASH_AGG;
  B.P.agg_nkw = A->nkw;
  B.P.agg_count = A->flags.as<uint64_t>();
  B.P.agg_overflow = (uint32_t*)(A->flags.as<uint64_t>() + 1);
  B.upload();
  const int64_t total = B.active_rows;
  auto launch_range = [&](int64_t b, int64_t e, uint64_t fill_limit) {
    B.P.begin = b;
    B.P.end = e;
    B.P.agg_slots = A->table.as<uint64_t>();
    B.P.agg_mask = A->cap - 1;
    B.P.agg_wrap = A->wrap;
    B.P.agg_fill_limit = fill_limit;
    B.P.agg_presized = fill_limit < A->cap ? 1 : 0;
    for (size_t s = 0; s < A->slots.size(); s++) B.P.acc_data[s] = A->acc[s].ptr;
    FlowLaunch F;
    // accumulator cache: as many entries (a power of two) as fit its budget,
    // 60 KB when the pipeline has no probes (nothing else is in its LDS, so a
    // few thousand hot groups stay on chip), else 16 KB
    bool probes = false;
    for (int st = 0; st < B.P.nsteps; st++) probes = probes || B.P.steps[st].kind == STEP_PROBE;
    const int per = 4 + 8 * (int)A->slots.size();
    const bool cacheable = (B.P.agg_mask + 1) <= 0xFFFFFFFEull;
    B.P.ac_n = 0;
    // the cache sits after the queues (the kernel keeps no cells for this sink);
    // its bytes are counted in the launch's LDS so the grid matches occupancy
    bool flow = false;
    for (const int budget : {probes ? 16384 : 61440, 16384}) {
      int acn = 4096;
      while (acn > 32 && acn * per > budget) acn >>= 1;
      if (!cacheable || acn * per > budget) break;
      if (plan_flow(B, SINK_HASH_AGG, sink_read_stages(B), (size_t)acn * per, F)) {
        flow = true;
        B.P.ac_n = acn;
        B.P.ac_off = 2 * B.P.qwords;  // u64 index past 4 waves x qwords u32
        break;
      }
    }
    if (!flow) flow = plan_flow(B, SINK_HASH_AGG, sink_read_stages(B), 0, F);
    if (flow) ctx->launch(B.P, F.grid, FLOW_BLOCK, F.lds);
    else ctx->launch(B.P, B.grid(), block_of(ctx), 0);
    A->len_known = false;
  };
  // under a memory budget the table grows (and spills) step by step
  const bool plain = A->sizing == ML_SIZING_PLAIN || ctx->mem->limit > 0;
  if (!plain && !A->cap && (uint64_t)A->hint >= PART_AGG_MIN_GROUPS && agg_partitioned(A, B)) return;
  if (plain) {
    // reserve(len + batch) before every device batch, doubling as needed
    const int64_t batch = ctx->cfg.batch_rows > 0 ? ctx->cfg.batch_rows : (1 << 21);
    for (int64_t b = 0; b < total || (b == 0 && total == 0); b += batch) {
      const int64_t e = std::min(total, b + batch);
      if (e <= b) break;
      if (!agg_reserve(A, agg_len(A) + (uint64_t)(e - b))) {
        // over budget (HashAggLocal::account, hash.rs:369-390): spill the
        // groups so far, then the batch must fit an empty table
        if (!ctx->mem->spill || A->no_spill) agg_exhausted(A, (uint64_t)(e - b));
        agg_spill(A);
        if (!agg_reserve(A, (uint64_t)(e - b))) agg_exhausted(A, (uint64_t)(e - b));
      }
      launch_range(b, e, A->cap);
    }
  } else {
    if (!A->cap) {
      A->len_host = 0;
      if (!agg_reserve(A, std::max<uint64_t>(1, (uint64_t)A->hint))) agg_exhausted(A, (uint64_t)A->hint);
    }
    launch_range(0, total, A->cap * 7 / 8);
  }
}

static uint32_t agg_overflow_bits(ml_hashagg* A) {
  uint64_t f[2] = {0, 0};
  d2h(A->ctx, f, A->flags.ptr, 16);
  A->len_host = f[0];
  A->len_known = true;
  return (uint32_t)f[1];
}

static ml_result* agg_finalize(ml_hashagg* A, bool states) {
  ml_ctx* ctx = A->ctx;
  const std::vector<OutDef>& outs = states ? A->state_outs : A->outs;
  auto* r = new ml_result();
  r->ctx = ctx;
  r->device = ctx->device;
  const size_t nk = A->key_types.size();
  if (!A->planned) {  // no input at all: empty result of the right shape
    // (without group keys, UngroupedAggregate's one row: COUNT 0, the others
    // NULL, ungrouped.rs:82-91)
    const int64_t rows = nk == 0 ? 1 : 0;
    for (size_t k = 0; k < nk; k++) {
      ResultColumn c;
      c.type = A->key_types[k];
      r->cols.push_back(std::move(c));
    }
    // Partial mode: the fresh states (COUNT 0; AVG sum 0.0 and count 0; the
    // others NULL, accumulator.rs:382-404); otherwise the finished values
    for (auto& d : A->user_aggs) {
      const bool is_count = d.func == ML_AGG_COUNT || d.func == ML_AGG_COUNT_STAR;
      std::vector<int> types{agg_final_type(d.func, d.arg_type)};
      if (A->mode == ML_AGG_MODE_PARTIAL) types = agg_state_types(d.func, d.arg_type);
      for (int t : types) {
        ResultColumn c;
        c.type = t;
        if (t == ML_TYPE_UTF8) {  // a NULL string over an empty string, as str_column writes it
          auto d = std::make_shared<HostDict>();
          d->entries.push_back("");
          d->finish_view();
          c.dict = d;
        }
        if (rows) {
          const uint64_t zero = 0;
          c.data = Buffer(ctx->arena, 8);
          upload(ctx, c.data.ptr, &zero, 8);
          if (!is_count && !(A->mode == ML_AGG_MODE_PARTIAL && d.func == ML_AGG_AVG)) {
            const uint8_t none = 0;
            c.valid = Buffer(ctx->arena, 1);
            upload(ctx, c.valid.ptr, &none, 1);
          }
        }
        r->cols.push_back(std::move(c));
      }
    }
    r->num_rows = rows;
    r->finish_views();
    return r;
  }
  if (A->direct) {
    const size_t ns = A->slots.size();
    std::vector<uint64_t> tot((size_t)A->ngroups * ns);
    d2h(ctx, tot.data(), A->totals.ptr, tot.size() * 8);
    std::vector<int64_t> groups;
    for (int64_t g = 0; g < A->ngroups; g++)  // without group keys the one group always (ungrouped.rs:75-96)
      if (tot[(size_t)g * ns] > 0 || nk == 0) groups.push_back(g);
    const size_t n = groups.size();
    r->num_rows = (int64_t)n;
    for (size_t k = 0; k < nk; k++) {
      ResultColumn c;
      c.type = A->key_types[k];
      c.dict = A->key_dicts[k];
      const int w = type_width(c.type);
      std::vector<uint8_t> host(std::max<size_t>(n, 1) * w);
      std::vector<uint8_t> valid(std::max<size_t>(n, 1), 1);
      bool any_null = false;
      for (size_t j = 0; j < n; j++) {
        int64_t code = (groups[j] / A->dim_stride[k]) % A->dim_size[k];
        const bool is_null = A->key_nullable && code == A->dim_size[k] - 1;
        if (is_null) { valid[j] = 0; any_null = true; code = 0; }
        if (w == 4) { uint32_t v = (uint32_t)code; std::memcpy(&host[j * 4], &v, 4); }
        else host[j] = (uint8_t)code;
      }
      c.data = Buffer(ctx->arena, host.size());
      upload(ctx, c.data.ptr, host.data(), host.size());
      if (any_null) {
        c.valid = Buffer(ctx->arena, valid.size());
        upload(ctx, c.valid.ptr, valid.data(), valid.size());
      }
      r->cols.push_back(std::move(c));
    }
    for (auto& o : outs) {
      ResultColumn c;
      c.type = o.out_type;
      if (o.finish == FIN_RANK) c.dict = A->ranks[(size_t)o.rank].dict;
      const int w = type_width(c.type);
      std::vector<uint8_t> host(std::max<size_t>(n, 1) * w);
      std::vector<uint8_t> valid(std::max<size_t>(n, 1), 1);
      bool any_null = false;
      for (size_t j = 0; j < n; j++) {
        const uint64_t* st = &tot[(size_t)groups[j] * ns];
        uint64_t v = st[o.state];
        const int64_t cnt = o.count >= 0 ? (int64_t)st[o.count] : 1;
        if (o.finish == FIN_AVG) {
          double s;
          std::memcpy(&s, &v, 8);
          const double avg = cnt == 0 ? 0.0 : s / (double)cnt;
          std::memcpy(&v, &avg, 8);
        } else if (o.finish == FIN_RANK) {
          const auto& hc = A->ranks[(size_t)o.rank].host_code;
          v = (cnt > 0 && v < hc.size()) ? hc[v] : 0;
        }
        if (o.count >= 0 && cnt <= 0) { valid[j] = 0; any_null = true; }
        if (w == 8) std::memcpy(&host[j * 8], &v, 8);
        else { uint32_t x = (uint32_t)v; std::memcpy(&host[j * 4], &x, 4); }
      }
      c.data = Buffer(ctx->arena, host.size());
      upload(ctx, c.data.ptr, host.data(), host.size());
      if (any_null) {
        c.valid = Buffer(ctx->arena, valid.size());
        upload(ctx, c.valid.ptr, valid.data(), valid.size());
      }
      r->cols.push_back(std::move(c));
    }
    r->finish_views();
    return r;
  }
  // hash mode: compact the occupied slots (ordered by slot) and decode
  if (agg_overflow_bits(A) & 2u) fail(ML_ERR_CAPACITY, "pre-sized group table overflowed");
  const uint64_t cap = A->cap;
  Buffer flags(ctx->arena, std::max<uint64_t>(cap, 1));
  launch_slot_flags(A->table.as<uint64_t>(), cap, A->nkw, flags.as<uint8_t>(), ctx->stream);
  const int64_t nb = compact_blocks((int64_t)cap);
  Buffer bc(ctx->arena, (size_t)std::max<int64_t>(nb, 1) * 4);
  Buffer offs(ctx->arena, (size_t)(nb + 1) * 8);
  Buffer idx(ctx->arena, std::max<uint64_t>(A->len_host, 1) * 8);
  launch_compact(flags.as<uint8_t>(), (int64_t)cap, bc.as<uint32_t>(), offs.as<uint64_t>(), nullptr, idx.ptr, 0,
                 ctx->stream);
  uint64_t n = 0;
  d2h(ctx, &n, offs.as<uint64_t>() + nb, 8);
  if (n != A->len_host) fail(ML_ERR_INTERNAL, "group count mismatch");
  r->num_rows = (int64_t)n;
  EmitPlan E{};
  E.nkeys = (int)nk;
  E.key_nullable = A->key_nullable ? 1 : 0;
  for (size_t k = 0; k < nk; k++)
    if (A->key_types[k] == ML_TYPE_FLOAT64) E.key_f64 |= 1 << k;
  for (size_t k = 0; k < nk; k++) {
    ResultColumn c;
    c.type = A->key_types[k];
    c.dict = A->key_dicts[k];
    const int w = type_width(c.type);
    c.data = Buffer(ctx->arena, std::max<uint64_t>(n, 1) * w);
    if (A->key_nullable) c.valid = Buffer(ctx->arena, std::max<uint64_t>(n, 1));
    E.key_width[k] = w;
    E.key_out[k] = c.data.ptr;
    E.key_valid[k] = c.valid.ptr ? c.valid.as<uint8_t>() : nullptr;
    r->cols.push_back(std::move(c));
  }
  E.nout = (int)outs.size();
  for (size_t a = 0; a < outs.size(); a++) {
    const OutDef& o = outs[a];
    ResultColumn c;
    c.type = o.out_type;
    const int w = type_width(c.type);
    c.data = Buffer(ctx->arena, std::max<uint64_t>(n, 1) * w);
    if (o.count > 0) c.valid = Buffer(ctx->arena, std::max<uint64_t>(n, 1));
    E.finish[a] = o.finish;
    E.out_width[a] = w;
    E.state[a] = A->acc[(size_t)o.state].ptr;
    E.count_state[a] = o.count >= 0 ? A->acc[(size_t)o.count].ptr : nullptr;
    E.out[a] = c.data.ptr;
    E.out_valid[a] = c.valid.ptr ? c.valid.as<uint8_t>() : nullptr;
    if (o.finish == FIN_RANK) {
      const RankMap& m = A->ranks[(size_t)o.rank];
      c.dict = m.dict;
      E.rank_code[a] = m.to_code.as<uint32_t>();
      E.rank_n[a] = m.host_code.size();
    }
    r->cols.push_back(std::move(c));
  }
  launch_agg_emit(A->table.as<uint64_t>(), A->nkw, idx.as<int64_t>(), (int64_t)n, E, ctx->stream);
  hip_check(hipStreamSynchronize(ctx->stream), "sync");
  r->finish_views();
  return r;
}

// Rows of several results of one schema, one after another.
static ml_result* concat_results(ml_ctx* ctx, std::vector<std::unique_ptr<ml_result>>& parts) {
  auto* r = new ml_result();
  r->ctx = ctx;
  r->device = ctx->device;
  int64_t total = 0;
  for (auto& p : parts) total += p->num_rows;
  r->num_rows = total;
  for (size_t c = 0; c < parts[0]->cols.size(); c++) {
    ResultColumn rc;
    rc.type = parts[0]->cols[c].type;
    rc.dict = parts[0]->cols[c].dict;
    const int w = type_width(rc.type);
    rc.data = Buffer(ctx->arena, (size_t)std::max<int64_t>(total, 1) * w);
    bool any_valid = false;
    for (auto& p : parts) any_valid = any_valid || p->cols[c].valid.ptr;
    if (any_valid) rc.valid = Buffer(ctx->arena, (size_t)std::max<int64_t>(total, 1));
    int64_t off = 0;
    for (auto& p : parts) {
      const int64_t n = p->num_rows;
      if (n > 0) {
        d2d(ctx, rc.data.as<uint8_t>() + off * w, p->cols[c].data.ptr, (size_t)n * w);
        if (any_valid) {
          if (p->cols[c].valid.ptr) {
            d2d(ctx, rc.valid.as<uint8_t>() + off, p->cols[c].valid.ptr, (size_t)n);
          } else {
            const std::vector<uint8_t> ones((size_t)n, 1);
            upload(ctx, rc.valid.as<uint8_t>() + off, ones.data(), ones.size());
          }
        }
      }
      off += n;
    }
    r->cols.push_back(std::move(rc));
  }
  r->finish_views();
  hip_check(hipStreamSynchronize(ctx->stream), "synchronize");
  return r;
}

// HashAggregateSink::finalize after a spill (hash.rs:394-425): the groups
// still in the table are spilled too; then each partition's states are merged
// by a Final aggregate of their own (about 1/16 of the groups in memory at a
// time, over budget now an error) and the partitions' rows concatenated. A
// Partial aggregate's merge emits states again.
static ml_result* agg_merge_spilled(ml_hashagg* A) {
  ml_ctx* ctx = A->ctx;
  agg_spill(A);
  const int nk = (int)A->key_types.size();
  std::vector<ml_aggregate> aggs(A->user_aggs.size());
  for (size_t a = 0; a < aggs.size(); a++) {
    aggs[a] = ml_aggregate{};
    aggs[a].func = A->user_aggs[a].func;
    aggs[a].arg_type = A->user_aggs[a].arg_type;
  }
  ml_agg_spec spec{};
  spec.key_types = A->key_types.data();
  spec.num_keys = nk;
  spec.num_aggs = (int32_t)aggs.size();
  spec.aggs = aggs.data();
  spec.mode = ML_AGG_MODE_FINAL;
  spec.sizing = ML_SIZING_PLAIN;
  auto check = [&](ml_status st) {
    if (st != ML_OK) fail(st, ctx->last_error);
  };
  std::vector<std::unique_ptr<ml_result>> parts;
  for (size_t p = 0; p < A->spill.size(); p++) {
    if (A->spill[p].empty()) continue;
    ml_hashagg* M = nullptr;
    check(ml_hashagg_create(ctx, &spec, &M));
    std::unique_ptr<ml_hashagg, void (*)(ml_hashagg*)> own(M, ml_hashagg_free);
    M->emit_states = A->mode == ML_AGG_MODE_PARTIAL || A->states_out;
    M->no_spill = true;
    for (auto& sb : A->spill[p]) {
      std::vector<Buffer> bufs;
      std::vector<ml_column> cols(sb.cols.size());
      for (size_t c = 0; c < sb.cols.size(); c++) {
        const auto& sc = sb.cols[c];
        Buffer d(ctx->arena, std::max<size_t>(sc.data.size(), 8));
        upload(ctx, d.ptr, sc.data.data(), sc.data.size());
        cols[c] = ml_column{};
        cols[c].type = sc.type;
        cols[c].data = d.ptr;
        cols[c].len = sb.rows;
        if (!sc.valid.empty()) {
          Buffer vb(ctx->arena, sc.valid.size());
          upload(ctx, vb.ptr, sc.valid.data(), sc.valid.size());
          Buffer bits(ctx->arena, (size_t)((sb.rows + 63) / 64) * 8);
          launch_bytes_to_bits(vb.as<uint8_t>(), sb.rows, bits.as<uint64_t>(), ctx->stream);
          cols[c].validity = bits.as<uint64_t>();
          bufs.push_back(std::move(vb));
          bufs.push_back(std::move(bits));
        }
        if ((int)c < nk && sc.type == ML_TYPE_UTF8) cols[c].dict = &A->key_dicts[c]->view;
        bufs.push_back(std::move(d));
      }
      ml_batch b{};
      b.columns = cols.data();
      b.num_columns = (int32_t)cols.size();
      b.num_rows = sb.rows;
      check(ml_hashagg_sink(M, &b));
      sb = ml_hashagg::SpillBatch{};  // the host copy is no longer needed
    }
    ml_result* r = nullptr;
    check(ml_hashagg_finalize(M, &r));
    parts.emplace_back(r);
  }
  A->spill.clear();
  if (parts.empty()) {  // nothing was ever held: an empty result of the aggregate's shape
    auto* r = new ml_result();
    r->ctx = ctx;
    r->device = ctx->device;
    for (int k = 0; k < nk; k++) {
      ResultColumn c;
      c.type = A->key_types[(size_t)k];
      c.dict = A->key_dicts.empty() ? nullptr : A->key_dicts[(size_t)k];
      r->cols.push_back(std::move(c));
    }
    for (auto& o : A->outs) {
      ResultColumn c;
      c.type = o.out_type;
      r->cols.push_back(std::move(c));
    }
    r->finish_views();
    return r;
  }
  return concat_results(ctx, parts);
}

static ml_result* agg_states(ml_hashagg* A);
static void agg_merge_states(ml_hashagg* G, ml_result* st);

}  // namespace mlg

// ---------------------------------------------------------------- join build
struct ml_join_build {
  ml_ctx* ctx = nullptr;
  std::mutex combine_mu;  // JoinBuildLocal::combine (join.rs:224) appends under the sink's lock
  std::vector<ml_expr> keys_raw;
  std::vector<int> key_types;
  std::vector<std::vector<ml_expr_node>> node_store;
  std::deque<std::string> str_store;
  bool schema_fixed = false;
  std::vector<int> col_types;
  std::vector<std::shared_ptr<mlg::HostDict>> col_dicts;
  std::vector<std::shared_ptr<mlg::HostDict>> key_dicts;
  std::vector<bool> col_nullable;
  int implied_col = -2;  // the build column equal to the first key (-1 none, -2 no batch yet)
  struct Chunk {
    std::vector<mlg::Buffer> cols, valid;
    mlg::Buffer keys, key_null;
    int64_t rows = 0;            // -1: still on the device, at count_dev
    int64_t cap = 0;             // rows the buffers hold
    mlg::Buffer count_buf;       // keeps count_dev alive
    const uint64_t* count_dev = nullptr;
    mlg::Buffer partials;        // per-workgroup key ranges of the kernel that wrote the rows, or empty
    int nparts = 0;
  };
  std::vector<Chunk> chunks;
};

namespace mlg {

constexpr uint64_t SMALL_BUILD_ROWS = 1 << 16;  // builds up to this many rows try the one-workgroup dense build

static void join_build_run(ml_join_build* J, PipelineBuild& B) {
  ml_ctx* ctx = J->ctx;
  if ((int)J->key_types.size() > MAX_KEYS) fail(ML_ERR_UNSUPPORTED, "more than 4 join keys");
  if ((int)B.layout.size() > MAX_COLS) fail(ML_ERR_UNSUPPORTED, "too many build columns");
  // schema of the build side
  std::vector<int> types;
  std::vector<std::shared_ptr<HostDict>> dicts;
  for (const VCol* v : B.layout) {
    types.push_back(v->type);
    dicts.push_back(v->dict);
  }
  std::vector<std::shared_ptr<HostDict>> kd;
  B.P.nkeys = (int)J->key_types.size();
  for (size_t k = 0; k < J->key_types.size(); k++) {
    NodePtr n = B.resolve(J->keys_raw[k]);
    std::shared_ptr<HostDict> d;
    if (J->key_types[k] == ML_TYPE_UTF8) {
      if (n->kind != ML_EXPR_COLUMN || !n->col->dict) fail(ML_ERR_UNSUPPORTED, "string join keys must be dictionary columns");
      d = n->col->dict;
    }
    kd.push_back(d);
    B.P.keys[k] = B.comp.compile_join_key(n, J->key_types[k]);
  }
  if (!J->schema_fixed) {
    J->col_types = types;
    J->col_dicts = dicts;
    J->key_dicts = kd;
    J->schema_fixed = true;
  } else {
    if (types != J->col_types) fail(ML_ERR_INTERNAL, "build batches with different schemas");
    for (size_t c = 0; c < dicts.size(); c++)
      if (!same_dict(dicts[c], J->col_dicts[c])) fail(ML_ERR_UNSUPPORTED, "build batches must share dictionaries");
  }
  uint64_t* counters = ctx->counters.as<uint64_t>();
  uint64_t cap = (uint64_t)B.active_rows;
  if (B.expanding) {
    B.P.sink = SINK_COUNT_PAIRS;
    B.P.out_count = counters;
    zero(ctx, counters, 16);
    B.upload();
    ctx->launch(B.P, B.grid(), block_of(ctx), 0);
    uint64_t n = 0;
    d2h(ctx, &n, counters, 8);
    cap = n;
  }
  ml_join_build::Chunk ch;
  B.P.ncols = (int)B.layout.size();
  std::vector<bool> nullable(B.layout.size());
  for (size_t c = 0; c < B.layout.size(); c++) {
    bool nl = false;
    B.P.cols[c] = column_prog(B, B.layout[c], &nl);
    nullable[c] = nl;
    const int w = type_width(B.layout[c]->type);
    B.P.col_width[c] = w;
    ch.cols.emplace_back(ctx->arena, std::max<uint64_t>(cap, 1) * w);
    ch.valid.emplace_back();
    if (nl) ch.valid.back() = Buffer(ctx->arena, std::max<uint64_t>(cap, 1));
    B.P.col_out[c] = ch.cols.back().ptr;
    B.P.col_valid[c] = ch.valid.back().ptr ? ch.valid.back().as<uint8_t>() : nullptr;
  }
  if (J->col_nullable.empty()) J->col_nullable = nullable;
  else
    for (size_t c = 0; c < nullable.size(); c++) J->col_nullable[c] = J->col_nullable[c] || nullable[c];
  // a column computed by exactly the first key's program (the key column
  // itself): a dense table held at key offsets need not store it
  int implied = -1;
  {
    const int kt = J->key_types[0];
    const Prog& kp = B.P.keys[0];
    if (kt == ML_TYPE_INT32 || kt == ML_TYPE_INT64 || kt == ML_TYPE_DATE)
      for (size_t c = 0; c < B.layout.size() && implied < 0; c++) {
        const Prog& cp = B.P.cols[c];
        if (B.layout[c]->type == kt && cp.count == kp.count && cp.nullable == kp.nullable && kp.count > 0 &&
            std::memcmp(&B.ins[(size_t)cp.first], &B.ins[(size_t)kp.first], sizeof(Instr) * (size_t)kp.count) == 0)
          implied = (int)c;
      }
  }
  J->implied_col = J->implied_col == -2 || J->implied_col == implied ? implied : -1;
  ch.keys = Buffer(ctx->arena, std::max<uint64_t>(cap, 1) * J->key_types.size() * 8);
  ch.key_null = Buffer(ctx->arena, std::max<uint64_t>(cap, 1));
  B.P.key_out = ch.keys.as<uint64_t>();
  B.P.key_null = ch.key_null.as<uint8_t>();
  B.P.sink = SINK_JOIN_BUILD;
  B.P.out_count = counters;
  B.P.out_capacity = cap;
  B.P.out_overflow = (uint32_t*)(counters + 1);
  B.upload();
  FlowLaunch F;
  ch.cap = (int64_t)cap;
  // no filter, probe or selection: every source row is a build row, written
  // at its own index by one row-per-lane kernel (no counting, no compaction)
  const bool identity = !B.expanding && B.P.nsteps == 0 && !B.P.sel;
  // a table too large for the one-workgroup build whose first key may make it
  // dense: the kernel that writes the rows also folds their first key words'
  // range per workgroup, so no separate pass over the keys finds it
  const int kt0 = J->key_types[0];
  const bool ranged = !B.expanding && cap > SMALL_BUILD_ROWS &&
                      (kt0 == ML_TYPE_INT32 || kt0 == ML_TYPE_INT64 || kt0 == ML_TYPE_DATE);
  if (!B.expanding && (cap <= SMALL_BUILD_ROWS || identity)) {
    // one append kernel, its count left on the device for the dense build's
    // first kernel (rows in any order: the table orders them)
    ch.count_buf = Buffer(ctx->arena, 16);
    if (identity && B.active_rows > 0) B.P.out_identity = 1;  // the kernel writes the count
    else zero(ctx, ch.count_buf.ptr, 16);
    B.P.out_count = ch.count_buf.as<uint64_t>();
    B.P.out_overflow = (uint32_t*)(ch.count_buf.as<uint64_t>() + 1);
    const int grid = B.grid();
    if (ranged) {
      ch.nparts = grid;
      ch.partials = Buffer(ctx->arena, (size_t)grid * 24);
      B.P.key_partials = ch.partials.as<int64_t>();
    }
    if (B.active_rows > 0) ctx->launch(B.P, grid, block_of(ctx), 0);
    B.P.out_identity = 0;
    B.P.key_partials = nullptr;
    ch.rows = -1;
    ch.count_dev = ch.count_buf.as<uint64_t>();
  } else if (plan_flow(B, SINK_ROWS, sink_read_stages(B), 0, F)) {
    if (ranged) {  // one partial per wave region (the materialization's workgroups)
      ch.nparts = (int)F.nwaves;
      ch.partials = Buffer(ctx->arena, (size_t)F.nwaves * 24);
      B.P.key_partials = ch.partials.as<int64_t>();
    }
    run_flow_rows(B, SINK_JOIN_BUILD, F, &ch.count_buf);  // the count is read with the key range
    B.P.key_partials = nullptr;
    ch.rows = -1;
    ch.count_dev = ch.count_buf.as<uint64_t>() + F.nwaves;
  } else {
    zero(ctx, counters, 16);
    ctx->launch(B.P, B.grid(), block_of(ctx), 0);
    uint64_t hc[2] = {0, 0};
    d2h(ctx, hc, counters, 16);
    if (hc[1] & 1) fail(ML_ERR_INTERNAL, "join build output overflowed");
    ch.rows = (int64_t)hc[0];
  }
  J->chunks.push_back(std::move(ch));
}

// Open-addressing table at load <= ML_JOIN_LOAD_PCT percent, built in LDS
// partitions when they fit (see join_part_build_kernel), else with global
// atomics; with a blocked Bloom filter from 64K rows (or as configured).
static uint64_t hashed_capacity(int64_t n) {
  return std::max<uint64_t>(16, next_pow2((uint64_t)((n * 100 + ML_JOIN_LOAD_PCT - 1) / ML_JOIN_LOAD_PCT)));
}
static bool hashed_bloom(ml_ctx* ctx, int64_t n) {
  const int use_bloom = ctx->cfg.use_bloom;
  return use_bloom > 0 || (use_bloom < 0 && n >= (1 << 16));
}
// one 64-bit block (3 bits set per key) per ML_BLOOM_KEYS_PER_WORD keys
static uint64_t hashed_bloom_words(int64_t n) {
  return std::max<uint64_t>(64, next_pow2((uint64_t)n / ML_BLOOM_KEYS_PER_WORD + 1));
}

// The global-atomic build (also the fallback when a partition overflows).
static void build_hashed_global(ml_ctx* ctx, JoinTableImpl& T, const Buffer& keys, const Buffer& key_null, int64_t n,
                                uint32_t* flags) {
  const uint64_t cap = T.dev.mask + 1;
  T.dev.wrap = cap - 1;
  T.dev.part_shift = __builtin_ctzll(cap);
  zero(ctx, flags, 8);
  zero(ctx, T.slots.ptr, T.slots.bytes);
  T.bloom = Buffer();
  T.dev.bloom = nullptr;
  if (hashed_bloom(ctx, n)) {
    const uint64_t bloom_words = hashed_bloom_words(n);
    T.bloom = Buffer(ctx->arena, bloom_words * 8);
    zero(ctx, T.bloom.ptr, T.bloom.bytes);
    T.dev.bloom = T.bloom.as<uint64_t>();
    T.dev.bloom_shift = __builtin_ctzll(bloom_words);
    T.dev.bloom_mask = bloom_words - 1;
  }
  launch_join_insert(T.dev, keys.as<uint64_t>(), key_null.as<uint8_t>(), n, T.bloom.as<uint64_t>(), ctx->stream);
  launch_join_unique(T.dev, flags, ctx->stream);
  uint32_t hdup = 0;
  d2h(ctx, &hdup, flags, 4);
  T.dev.unique = hdup ? 0 : 1;
}

// `defer`: leave the partitioned build's check on the stream (the table is
// then pending and keeps `keys`/`key_null` for a fallback).
static void build_hashed(ml_ctx* ctx, JoinTableImpl& T, Buffer& keys, Buffer& key_null, int64_t n, size_t nk,
                         uint32_t* flags, bool defer) {
  const uint64_t cap = hashed_capacity(n);
  T.dev.mask = cap - 1;
  T.slots = Buffer(ctx->arena, cap * (1 + nk) * 8);
  T.dev.slots = T.slots.as<uint64_t>();
  const bool bloom = hashed_bloom(ctx, n);
  const uint64_t bloom_words = bloom ? hashed_bloom_words(n) : 0;
  zero(ctx, flags, 8);
  // Partitioned build: partitions of S slots, each built by one workgroup in
  // LDS (no global atomics), S chosen so a partition's slots fit in 32 KB.
  const uint64_t smax = nk == 1 ? 2048 : (nk <= 3 ? 1024 : 512);
  const uint64_t S = std::min(cap, smax);
  const uint64_t nparts = cap / S;
  if (n > 0 && nparts <= 8192 && (uint64_t)n < 0xFFFFFFFFull) {
    T.dev.wrap = S - 1;
    T.dev.part_shift = __builtin_ctzll(S);
    uint64_t W = 0;
    if (bloom) {
      W = std::min<uint64_t>(2048, std::max<uint64_t>(1, bloom_words / nparts));
      T.bloom = Buffer(ctx->arena, W * nparts * 8);
      T.dev.bloom = T.bloom.as<uint64_t>();
      T.dev.bloom_shift = __builtin_ctzll(W);
      T.dev.bloom_mask = W - 1;
    }
    // workgroups of >= 2048 rows (8 per thread), at most 1024 of them
    const int nblocks = (int)std::min<int64_t>(1024, std::max<int64_t>(1, (n + 2047) / 2048));
    const int64_t rpb = (n + nblocks - 1) / nblocks;
    const int64_t nh = (int64_t)nparts * nblocks;
    Buffer hist(ctx->arena, (size_t)nh * 4), offsets(ctx->arena, (size_t)(nh + 1) * 4);
    Buffer sums(ctx->arena, (size_t)(scan_u32_blocks(nh) + 1) * 4);
    Buffer srow(ctx->arena, (size_t)n * 4), skeys(ctx->arena, (size_t)n * nk * 8);
    launch_join_part_hist(T.dev, keys.as<uint64_t>(), key_null.as<uint8_t>(), n, nblocks, rpb, (int)nparts,
                          hist.as<uint32_t>(), ctx->stream);
    launch_scan_u32(hist.as<uint32_t>(), nh, offsets.as<uint32_t>(), sums.as<uint32_t>(), ctx->stream);
    launch_join_part_scatter(T.dev, keys.as<uint64_t>(), key_null.as<uint8_t>(), n, nblocks, rpb, (int)nparts,
                             offsets.as<uint32_t>(), srow.as<uint32_t>(), skeys.as<uint64_t>(), ctx->stream);
    launch_join_part_build(T.dev, (int)nparts, offsets.as<uint32_t>(), nblocks, srow.as<uint32_t>(),
                           skeys.as<uint64_t>(), T.dev.bloom ? T.bloom.as<uint64_t>() : nullptr, flags,
                           S * (1 + nk) * 8 + W * 8, ctx->stream);
    if (defer && defer_check(ctx, T, flags, PENDING_PARTS)) {
      T.dev.unique = 1;  // settled by join_table_resolve
      T.keys = std::move(keys);
      T.key_null = std::move(key_null);
      return;
    }
    uint32_t hf = 0;
    d2h(ctx, &hf, flags, 4);
    if (!(hf & 2)) {
      T.dev.unique = (hf & 1) ? 0 : 1;
      return;
    }
    // a partition overflowed (many equal keys): build unpartitioned
  }
  build_hashed_global(ctx, T, keys, key_null, n, flags);
}

static void refresh_views(JoinTableImpl& T) {
  T.views.resize(T.col_types.size());
  for (size_t c = 0; c < T.col_types.size(); c++) {
    ml_column v{};
    v.type = T.col_types[c];
    // an offset-indexed table's views are its compacted copies (made on first access)
    v.data = T.dense_offset ? (c < T.compact_cols.size() ? T.compact_cols[c].ptr : nullptr) : T.cols[c].ptr;
    v.len = T.nrows;
    v.dict = T.col_dicts[c] ? &T.col_dicts[c]->view : nullptr;
    T.views[c] = v;
  }
}

// Second key words that all fit a byte (flag bit 8 clear): probes read the
// byte copy (a quarter of the u32 words' footprint in L2).
static void use_keys8(JoinTableImpl& T, uint32_t hf) {
  if (T.dense_keys8.ptr && !(hf & 256u)) {
    T.dense_keys = std::move(T.dense_keys8);
    T.dev.dense_keys = T.dense_keys.as<uint64_t>();
    T.dev.dense_keys32 = 0;
    T.dev.dense_keys8 = 1;
  }
  T.dense_keys8 = Buffer();
}

void join_table_resolve(ml_join_table* t) {
  JoinTableImpl& T = t->impl;
  if (!T.pending) return;
  ml_ctx* ctx = t->ctx;
  // wait for this check's own copy: the kernels queued behind it keep the
  // device busy while the host plans with the settled layout
  if (T.check_seq > ctx->synced_seq) hip_check(hipEventSynchronize(T.checks->event(T.check)), "synchronize");
  const uint32_t hf = T.checks->words[T.check];
  T.checks->give(T.check);
  T.check = -1;
  const int kind = T.pending;
  T.pending = PENDING_NONE;
  uint32_t* flags = (uint32_t*)ctx->counters.as<uint64_t>();
  const size_t nk = T.key_types.size();
  if (kind == PENDING_DENSE && (hf & 5)) {
    // a repeated key (or one outside the range): back to build order and the hashed layout
    T.cols = std::move(T.orig_cols);
    T.col_valid = std::move(T.orig_valid);
    T.dense_bits = Buffer();
    T.dense_rank = Buffer();
    T.dense_keys = Buffer();
    T.dense_keys8 = Buffer();
    T.bloom = Buffer();
    T.dense_offset = false;
    T.offset_rows = 0;
    T.implied_col = -1;
    T.dev.dense = 0;
    T.dev.dense_full = 0;
    T.dev.dense_keys8 = 0;
    T.dev.dense_bits = nullptr;
    T.dev.dense_rank = nullptr;
    T.dev.dense_keys = nullptr;
    T.dev.dense_keys32 = 0;
    T.dev.bloom = nullptr;
    refresh_views(T);
    build_hashed(ctx, T, T.keys, T.key_null, T.nrows, nk, flags, false);
  } else if (kind == PENDING_DENSE) {
    use_keys8(T, hf);
  } else if (kind == PENDING_PARTS) {
    if (hf & 2) build_hashed_global(ctx, T, T.keys, T.key_null, T.nrows, flags);
    else T.dev.unique = (hf & 1) ? 0 : 1;
  }
  T.keys = Buffer();
  T.key_null = Buffer();
  T.orig_cols.clear();
  T.orig_valid.clear();
}

// A join build's finalize runs in two parts around one read of the device:
// the first enqueues what the host must know before it chooses the layout
// (a small table's whole dense build and its status, or a large table's row
// count and first-key range) with a copy into a pinned slot; the second reads
// the slot and finishes. Several builds can share one synchronization.
struct FinalizePrep {
  bool small = false;   // the one-workgroup dense build ran; its status is in the slot
  bool ranged = false;  // key_minmax ran; count and range are in the slot
  bool batched = false; // other builds' kernels run between the two parts
  int64_t slot[4] = {0, -1, 0, 2};
  DensePermute M{};
  std::vector<Buffer> cols, valid;
  Buffer dkeys, bits, rank, status, partials, mm;
};

static bool finalize_keys32(const std::vector<int>& key_types, uint32_t* sx) {
  bool keys32 = key_types.size() > 1;
  *sx = 0;
  for (size_t k = 1; k < key_types.size(); k++) {
    const int kt = key_types[k];
    keys32 = keys32 && (kt == ML_TYPE_INT32 || kt == ML_TYPE_DATE || kt == ML_TYPE_UTF8);
    if (kt == ML_TYPE_INT32 || kt == ML_TYPE_DATE) *sx |= 1u << k;
  }
  return keys32;
}

// First part: enqueue the build's device read into `pinned` (4 x i64).
static void join_finalize_enqueue(ml_join_build* J, FinalizePrep& P, int64_t* pinned) {
  ml_ctx* ctx = J->ctx;
  const size_t nk = J->key_types.size();
  uint32_t* flags = (uint32_t*)ctx->counters.as<uint64_t>();
  const int kt0 = J->key_types[0];
  const bool dense_type = kt0 == ML_TYPE_INT32 || kt0 == ML_TYPE_INT64 || kt0 == ML_TYPE_DATE;
  uint32_t keys_sx;
  const bool keys32 = finalize_keys32(J->key_types, &keys_sx);
  const size_t key_word_bytes = keys32 ? 4 : 8;
  if (J->chunks.size() == 1 && J->chunks[0].rows < 0 && dense_type && (uint64_t)J->chunks[0].cap <= SMALL_BUILD_ROWS) {
    // small table: the whole dense build in one workgroup, one read of its status
    auto& c = J->chunks[0];
    const int64_t cap = std::max<int64_t>(c.cap, 1);
    DensePermute& M = P.M;
    if (nk > 1) P.dkeys = Buffer(ctx->arena, (size_t)cap * (nk - 1) * key_word_bytes);
    M.nkw = (int32_t)nk;
    M.keys32 = keys32 ? 1 : 0;
    M.dst_keys = P.dkeys.as<uint64_t>();
    M.ncols = (int32_t)c.cols.size();
    for (size_t ci = 0; ci < c.cols.size(); ci++) {
      const int w = type_width(J->col_types[ci]);
      P.cols.emplace_back(ctx->arena, (size_t)cap * w);
      P.valid.emplace_back();
      const bool has_valid = ci < c.valid.size() && c.valid[ci].ptr;
      if (has_valid) P.valid.back() = Buffer(ctx->arena, (size_t)cap);
      M.width[ci] = w;
      M.src[ci] = c.cols[ci].ptr;
      M.dst[ci] = P.cols.back().ptr;
      M.src_valid[ci] = has_valid ? c.valid[ci].as<uint8_t>() : nullptr;
      M.dst_valid[ci] = has_valid ? P.valid.back().as<uint8_t>() : nullptr;
    }
    const int words = small_dense_words();
    P.bits = Buffer(ctx->arena, (size_t)words * 8);
    P.rank = Buffer(ctx->arena, (size_t)(words + 1) * 4);
    P.status = Buffer(ctx->arena, 32);
    launch_small_dense(M, c.keys.as<uint64_t>(), c.key_null.as<uint8_t>(), c.cap, c.count_dev, P.bits.as<uint64_t>(),
                       P.rank.as<uint32_t>(), P.status.as<int64_t>(), ctx->stream);
    hip_check(hipMemcpyAsync(pinned, P.status.ptr, 32, hipMemcpyDeviceToHost, ctx->stream), "copy to host");
    P.small = true;
    return;
  }
  if (J->chunks.size() == 1 && J->chunks[0].rows < 0 && dense_type && (uint64_t)J->chunks[0].cap < 0xFFFFFFFFull) {
    auto& c = J->chunks[0];
    P.mm = Buffer(ctx->arena, 32);
    // (both also clear flags[0] for the presence marking and flags[1] for the permutation)
    if (c.partials.ptr) {
      // the rows' kernel left per-workgroup ranges: one workgroup folds them
      launch_key_range_reduce(c.partials.as<int64_t>(), c.nparts, c.cap, c.count_dev, P.mm.as<int64_t>(), flags,
                              ctx->stream);
    } else {
      // at most 256 workgroups: each ends with one atomic on a shared counter
      const int nparts = (int)std::min<int64_t>(256, std::max<int64_t>(1, (c.cap + 4095) / 4096));
      P.partials = Buffer(ctx->arena, (size_t)nparts * 24);
      launch_key_minmax(c.keys.as<uint64_t>(), (int)nk, c.key_null.as<uint8_t>(), c.cap, c.count_dev,
                        P.partials.as<int64_t>(), nparts, P.mm.as<int64_t>(),
                        (uint32_t*)(ctx->counters.as<uint64_t>() + ml_ctx::DONE_WORD), flags, ctx->stream);
    }
    hip_check(hipMemcpyAsync(pinned, P.mm.ptr, 32, hipMemcpyDeviceToHost, ctx->stream), "copy to host");
    P.ranged = true;
  }
}

// Second part, after the stream has passed the copy and P.slot holds it.
static ml_join_table* join_finalize_finish(ml_join_build* J, FinalizePrep& P) {
  ml_ctx* ctx = J->ctx;
  auto* t = new ml_join_table();
  t->ctx = ctx;
  JoinTableImpl& T = t->impl;
  T.key_types = J->key_types;
  T.key_dicts = J->key_dicts;
  T.col_types = J->col_types;
  T.col_dicts = J->col_dicts;
  if (T.key_dicts.size() < T.key_types.size()) T.key_dicts.resize(T.key_types.size());
  const size_t nk = J->key_types.size();
  uint32_t* flags = (uint32_t*)ctx->counters.as<uint64_t>();
  const int kt0 = T.key_types[0];
  const bool dense_type = kt0 == ML_TYPE_INT32 || kt0 == ML_TYPE_INT64 || kt0 == ML_TYPE_DATE;
  uint32_t keys_sx;
  const bool keys32 = finalize_keys32(T.key_types, &keys_sx);
  const size_t key_word_bytes = keys32 ? 4 : 8;
  // the later dense steps use flags[0..1] from zero (key_minmax cleared them,
  // unless other builds ran in between)
  if (P.batched) launch_fill_u64(ctx->counters.as<uint64_t>(), 0, 1, ctx->stream);
  int64_t hm[4] = {0, -1, 0, 0};  // min, max, rows, non-NULL rows
  bool have_range = false;
  auto key_range = [&](const uint64_t* kw, const uint8_t* kn, int64_t rows_cap, const uint64_t* count_dev) {
    const int nparts = (int)std::min<int64_t>(256, std::max<int64_t>(1, (rows_cap + 4095) / 4096));
    P.partials = Buffer(ctx->arena, (size_t)nparts * 24);
    P.mm = Buffer(ctx->arena, 32);
    launch_key_minmax(kw, (int)nk, kn, rows_cap, count_dev, P.partials.as<int64_t>(), nparts, P.mm.as<int64_t>(),
                      (uint32_t*)(ctx->counters.as<uint64_t>() + ml_ctx::DONE_WORD), flags, ctx->stream);
    d2h(ctx, hm, P.mm.ptr, 32);
    have_range = true;
  };
  if (P.small) {
    auto& c = J->chunks[0];
    const int64_t* st = P.slot;
    c.rows = st[0];
    if ((st[3] & 3) == 0) {
      T.nrows = st[0];
      T.cols = std::move(P.cols);
      T.col_valid = std::move(P.valid);
      T.dense_bits = std::move(P.bits);
      T.dense_rank = std::move(P.rank);
      T.dense_keys = std::move(P.dkeys);
      T.dev.nkw = (int)nk;
      T.dev.dense = 1;
      T.dev.dense_full = (st[3] & SMALL_DENSE_FULL) ? 1 : 0;
      T.dev.dense_min = st[1];
      T.dev.dense_range = (uint64_t)st[2];
      T.dev.dense_bits = T.dense_bits.as<uint64_t>();
      T.dev.dense_rank = T.dense_rank.as<uint32_t>();
      T.dev.dense_keys = T.dense_keys.as<uint64_t>();
      T.dev.dense_keys32 = keys32 ? 1 : 0;
      T.dev.dense_keys_sx = keys_sx;
      T.dev.unique = 1;
      J->chunks.clear();
      refresh_views(T);
      return t;
    }
  } else if (P.ranged) {
    std::memcpy(hm, P.slot, 32);
    have_range = true;
    J->chunks[0].rows = hm[2];
  }
  for (auto& c : J->chunks)
    if (c.rows < 0) d2h(ctx, &c.rows, c.count_dev, 8);
  int64_t n = 0;
  for (auto& c : J->chunks) n += c.rows;
  T.nrows = n;
  Buffer keys, key_null;
  if (J->chunks.size() == 1) {
    auto& c = J->chunks[0];
    T.cols = std::move(c.cols);
    T.col_valid = std::move(c.valid);
    keys = std::move(c.keys);
    key_null = std::move(c.key_null);
  } else {
    // Batch::concat of the collected batches (batch.rs:168-184)
    for (size_t ci = 0; ci < T.col_types.size(); ci++) {
      const int w = type_width(T.col_types[ci]);
      T.cols.emplace_back(ctx->arena, (size_t)std::max<int64_t>(n, 1) * w);
      const bool nl = ci < J->col_nullable.size() && J->col_nullable[ci];
      T.col_valid.emplace_back();
      if (nl) {
        T.col_valid.back() = Buffer(ctx->arena, (size_t)std::max<int64_t>(n, 1));
        launch_fill_u64(T.col_valid.back().as<uint64_t>(), 0x0101010101010101ull,
                        (std::max<int64_t>(n, 1) + 7) / 8, ctx->stream);
      }
      int64_t off = 0;
      for (auto& c : J->chunks) {
        d2d(ctx, T.cols[ci].as<uint8_t>() + off * w, c.cols[ci].ptr, (size_t)c.rows * w);
        if (nl && c.valid[ci].ptr) d2d(ctx, T.col_valid[ci].as<uint8_t>() + off, c.valid[ci].ptr, (size_t)c.rows);
        off += c.rows;
      }
    }
    keys = Buffer(ctx->arena, (size_t)std::max<int64_t>(n, 1) * nk * 8);
    key_null = Buffer(ctx->arena, (size_t)std::max<int64_t>(n, 1));
    int64_t off = 0;
    for (auto& c : J->chunks) {
      d2d(ctx, keys.as<uint64_t>() + off * nk, c.keys.ptr, (size_t)c.rows * nk * 8);
      d2d(ctx, key_null.as<uint8_t>() + off, c.key_null.ptr, (size_t)c.rows);
      off += c.rows;
    }
  }
  J->chunks.clear();
  if (T.col_valid.size() < T.cols.size()) T.col_valid.resize(T.cols.size());
  T.dev.nkw = (int)nk;
  // Dense layout: a first key column of integers whose values span at most
  // max(8n, 2^24) consecutive integers, all distinct (at the floor the
  // presence bits are 2 MB, held in L2 like a Bloom filter). The rows move into
  // first-key order; a probe reads one presence bit and, for a present key,
  // the rank of its 64-key word: no hashing, probing or row index array.
  // With several key columns the other key words are compared per row, and a
  // Bloom filter over the whole key rejects most misses first (the first
  // column alone would not reject rows the full key does).
  if (n > 0 && (uint64_t)n < 0xFFFFFFFFull && dense_type) {
    if (!have_range) key_range(keys.as<uint64_t>(), key_null.as<uint8_t>(), n, nullptr);
    if (hm[0] <= hm[1]) {
      const unsigned __int128 range = (unsigned __int128)((__int128)hm[1] - (__int128)hm[0]) + 1;
      const unsigned __int128 limit = std::max<uint64_t>((uint64_t)n * 8, 1ull << 24);
      if (range <= limit && range < ((unsigned __int128)1 << 31)) {
        const uint64_t R = (uint64_t)range;
        const int64_t ne = (int64_t)((R + 63) / 64);
        Buffer bits(ctx->arena, (size_t)ne * 8);
        zero(ctx, bits.ptr, bits.bytes);
        // every key of the range present (all distinct, or the table is
        // rebuilt hashed): a row's rank is its key's offset, so no ranks are
        // computed, and a probe reads neither bits nor ranks; a Bloom filter
        // would cost the one load it saves
        const bool full = hm[3] == (int64_t)R;
        // otherwise the rows are held at their key offset when that costs at
        // most 4x the rows' bytes or 256 MB: a probe then reads the presence
        // bit and the row at the key's offset, with no rank directory
        // between them (one dependent load fewer per hit)
        const int64_t nnull = n - hm[3];
        const uint64_t offset_rows = R + (uint64_t)nnull;
        uint64_t row_bytes = (nk - 1) * key_word_bytes + (nk == 2 ? 1 : 0);
        for (size_t c = 0; c < T.cols.size(); c++)
          if ((int)c != J->implied_col)  // not stored when the rows sit at key offsets
            row_bytes += (uint64_t)type_width(T.col_types[c]) + (T.col_valid[c].ptr ? 1 : 0);
        const bool offset = !full && (offset_rows <= 4 * (uint64_t)n || offset_rows * row_bytes <= (256ull << 20));
        const int64_t held = (full || offset) ? (int64_t)offset_rows : n;  // full: R + NULL-key rows = n
        const int implied = (full || offset) ? J->implied_col : -1;
        DensePermute M{};
        Buffer dkeys, dkeys8;
        std::vector<Buffer> cols, valid;
        auto prepare_rows = [&] {
          if (nk > 1) dkeys = Buffer(ctx->arena, (size_t)std::max<int64_t>(held, 1) * (nk - 1) * key_word_bytes);
          if (nk == 2) {  // also a byte per row, used if every second key word fits (flag bit 8)
            dkeys8 = Buffer(ctx->arena, (size_t)std::max<int64_t>(held, 1));
            M.dst_keys8 = dkeys8.as<uint8_t>();
          }
          M.nkw = (int32_t)nk;
          M.keys32 = keys32 ? 1 : 0;
          M.dst_keys = dkeys.as<uint64_t>();
          M.held = held;
          M.ncols = (int32_t)T.cols.size();
          for (size_t c = 0; c < T.cols.size(); c++) {
            const int w = type_width(T.col_types[c]);
            M.width[c] = w;
            if ((int)c == implied) {  // not written (dst null): row p holds key min + p
              cols.emplace_back();
              valid.emplace_back();
              M.src[c] = M.dst[c] = nullptr;
              M.src_valid[c] = M.dst_valid[c] = nullptr;
              continue;
            }
            cols.emplace_back(ctx->arena, (size_t)std::max<int64_t>(held, 1) * w);
            valid.emplace_back();
            if (T.col_valid[c].ptr) valid.back() = Buffer(ctx->arena, (size_t)std::max<int64_t>(held, 1));
            M.src[c] = T.cols[c].ptr;
            M.dst[c] = cols.back().ptr;
            M.src_valid[c] = T.col_valid[c].ptr ? T.col_valid[c].as<uint8_t>() : nullptr;
            M.dst_valid[c] = valid.back().ptr ? valid.back().as<uint8_t>() : nullptr;
          }
        };
        // rows at key offsets: marking and moves in one pass (flags[0..1]
        // were cleared by the range kernel). Ranked: the marking, then the
        // ranks and the moves once the repeated-key flag allows them.
        const bool place = full || offset;
        if (place) {
          prepare_rows();
          launch_dense_place(M, keys.as<uint64_t>(), key_null.as<uint8_t>(), n, hm[0], R, bits.as<uint64_t>(), flags,
                             flags + 1, ctx->stream);
        } else {
          launch_dense_mark(keys.as<uint64_t>(), (int)nk, key_null.as<uint8_t>(), n, hm[0], R, bits.as<uint64_t>(),
                            flags, ctx->stream);
        }
        // the repeated-key flag is read when the table is first used; the
        // rank and permutation run meanwhile (a repeated key discards them)
        const bool deferred = defer_check(ctx, T, flags, PENDING_DENSE);
        uint32_t hf = 0;
        if (!deferred) d2h(ctx, &hf, flags, 4);
        if (!(hf & 5)) {
          Buffer rank;
          if (!place) {
            rank = Buffer(ctx->arena, (size_t)(ne + 1) * 4);
            Buffer sums(ctx->arena, (size_t)(scan_u32_blocks(ne) + 1) * 4);
            launch_dense_rank(bits.as<uint64_t>(), ne, rank.as<uint32_t>(), sums.as<uint32_t>(), ctx->stream);
            prepare_rows();
            // flags[1] (next NULL-key row) was cleared by the range kernel
            launch_dense_permute(M, keys.as<uint64_t>(), key_null.as<uint8_t>(), n, hm[0], R, bits.as<uint64_t>(),
                                 rank.as<uint32_t>(), rank.as<uint32_t>() + ne, flags + 1, ctx->stream);
          }
          if (nk > 1 && !full && hashed_bloom(ctx, n)) {  // a word per 8 first-key values
            const uint64_t words = (R + (1u << DENSE_BLOOM_SHIFT) - 1) >> DENSE_BLOOM_SHIFT;
            T.bloom = Buffer(ctx->arena, words * 8);
            zero(ctx, T.bloom.ptr, T.bloom.bytes);
            T.dev.bloom = T.bloom.as<uint64_t>();
            launch_dense_bloom(T.dev, keys.as<uint64_t>(), key_null.as<uint8_t>(), n, hm[0], R,
                               T.bloom.as<uint64_t>(), ctx->stream);
          }
          if (deferred) {
            T.orig_cols = std::move(T.cols);
            T.orig_valid = std::move(T.col_valid);
            T.keys = std::move(keys);
            T.key_null = std::move(key_null);
          }
          T.cols = std::move(cols);
          T.col_valid = std::move(valid);
          T.dense_bits = std::move(bits);
          T.dense_rank = std::move(rank);
          T.dense_keys = std::move(dkeys);
          T.dev.dense_keys = T.dense_keys.as<uint64_t>();
          T.dev.dense_keys32 = keys32 ? 1 : 0;
          T.dev.dense_keys_sx = keys_sx;
          T.dev.dense = 1;
          T.dev.dense_full = full ? 1 : 0;
          T.dev.dense_min = hm[0];
          T.dev.dense_range = R;
          T.dev.dense_bits = T.dense_bits.as<uint64_t>();
          T.dev.dense_rank = T.dense_rank.ptr ? T.dense_rank.as<uint32_t>() : nullptr;
          T.dev.unique = 1;
          T.dense_offset = offset;
          T.offset_rows = offset ? (int64_t)offset_rows : 0;
          T.implied_col = implied;
          T.dense_keys8 = std::move(dkeys8);
          if (!deferred) use_keys8(T, hf);
        }
      }
    }
  }
  if (!T.dev.dense) build_hashed(ctx, T, keys, key_null, n, nk, flags, true);
  refresh_views(T);
  return t;
}

}  // namespace mlg

// -------------------------------------------------------------------- TopN
struct ml_topn {
  ml_ctx* ctx = nullptr;
  std::vector<ml_expr> keys_raw;
  std::vector<int> asc, nulls_first;
  int64_t limit = 0;
  std::vector<std::vector<ml_expr_node>> node_store;
  std::deque<std::string> str_store;
  bool schema_fixed = false;
  mlg::Materialized best;  // the best `limit` rows seen so far
  std::mutex combine_mu;   // SortLocal::combine (sort.rs:253-268) merges under the sink's lock
};

namespace mlg {

// Sort a materialized table's rows by the TopN keys and keep the first `limit`.
static void topn_reduce(ml_topn* T, Materialized& m) {
  ml_ctx* ctx = T->ctx;
  const int64_t n = m.rows;
  const int64_t L = T->limit;  // < 0: no limit (SortSink without LIMIT)
  if (n == 0 || L == 0) {
    m.rows = 0;
    return;
  }
  // layout over the materialized columns
  PipelineBuild B(ctx);
  for (size_t c = 0; c < m.types.size(); c++) {
    VCol v;
    v.type = m.types[c];
    v.data = m.cols[c].ptr;
    v.valid = m.valid[c].ptr ? m.valid[c].as<uint8_t>() : nullptr;
    v.dict = m.dicts[c];
    B.layout.push_back(B.add_col(std::move(v)));
  }
  SortKeys K{};
  K.nkeys = (int)T->keys_raw.size();
  if (K.nkeys > MAX_SORT_KEYS) fail(ML_ERR_UNSUPPORTED, "more than 4 sort keys");
  int nw = 0;
  for (int k = 0; k < K.nkeys; k++) {
    NodePtr e = B.resolve(T->keys_raw[(size_t)k]);
    const int t = B.comp.type_of(e);
    Prog p;
    if (t == ML_TYPE_UTF8) {
      // strings sort bytewise: rank of each dictionary entry
      if (e->kind != ML_EXPR_COLUMN || !e->col->dict) fail(ML_ERR_UNSUPPORTED, "string sort keys must be dictionary columns");
      const auto& ent = e->col->dict->entries;
      std::vector<uint32_t> order(ent.size());
      std::iota(order.begin(), order.end(), 0u);
      std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return ent[a] < ent[b]; });
      std::vector<uint64_t> rank(ent.size() + 1, 0);
      for (size_t r = 0; r < order.size(); r++) {
        uint64_t rk = r;
        if (r > 0 && ent[order[r]] == ent[order[r - 1]]) rk = rank[order[r - 1]];
        rank[order[r]] = rk;
      }
      Buffer tb(ctx->arena, rank.size() * 8);
      upload(ctx, tb.ptr, rank.data(), rank.size() * 8);
      p = B.comp.compile(e, -1);
      Instr in{};
      in.op = OP_LOOKUP_U64;
      in.depth = 1;
      in.imm = (uint64_t)(uintptr_t)tb.as<uint64_t>();
      B.ins.push_back(in);
      p.count++;
      B.tables.push_back(std::move(tb));
      K.kind[k] = SK_UNSIGNED;
    } else {
      p = B.comp.compile(e, -1);
      K.kind[k] = t == ML_TYPE_FLOAT64 ? SK_F64 : (t == ML_TYPE_BOOLEAN ? SK_UNSIGNED : SK_INT);
    }
    K.prog[k] = p;
    K.asc[k] = T->asc[(size_t)k];
    K.nulls_first[k] = T->nulls_first[(size_t)k];
    K.nullable[k] = p.nullable;
    nw += p.nullable ? 2 : 1;
  }
  K.nwords = nw;
  if (nw + 1 > 6) fail(ML_ERR_UNSUPPORTED, "sort keys too wide");
  B.upload();
  const int W = nw + 1;
  Buffer words(ctx->arena, (size_t)n * W * 8);
  launch_sort_keys(B.P, K, n, words.as<uint64_t>(), ctx->stream);
  const int64_t m_out = L < 0 ? n : std::min<int64_t>(L, n);
  Buffer rows(ctx->arena, (size_t)m_out * 8);
  if (L < 0 || L > 256) {
    // a full sort (sort_batches without or with a large limit, sort.rs:85-132)
    const int64_t N = sort_padded_entries(n);
    Buffer s1(ctx->arena, (size_t)N * W * 8), s2(ctx->arena, (size_t)N * W * 8);
    Buffer split(ctx->arena, (size_t)sort_split_entries(n) * 8);
    const uint64_t* sorted = launch_sort_words(words.as<uint64_t>(), n, nw, s1.as<uint64_t>(), s2.as<uint64_t>(),
                                               split.as<int64_t>(), ctx->stream);
    launch_topn_rows(sorted, nw, m_out, rows.as<int64_t>(), ctx->stream);
  } else {
    // repeated tile selection until one tile remains: rank selection over
    // 256-entry tiles for a small limit, sorted 1024-entry tiles otherwise
    const bool by_rank = L <= TOPN_RANK_LIMIT;
    const int tile = by_rank ? TOPN_RANK_TILE : TOPN_TILE;
    int64_t cur = n;
    Buffer a = std::move(words);
    for (;;) {
      const int64_t tiles = (cur + tile - 1) / tile;
      const int keep = (int)std::min<int64_t>(L, tile);
      Buffer b(ctx->arena, (size_t)tiles * keep * W * 8);
      if (by_rank) launch_topn_rank(a.as<uint64_t>(), cur, nw, keep, b.as<uint64_t>(), ctx->stream);
      else launch_topn_tiles(a.as<uint64_t>(), cur, nw, keep, b.as<uint64_t>(), ctx->stream);
      a = std::move(b);
      cur = tiles * keep;
      if (tiles == 1) break;
    }
    launch_topn_rows(a.as<uint64_t>(), nw, m_out, rows.as<int64_t>(), ctx->stream);
  }
  // gather every column of the winners
  for (size_t c = 0; c < m.types.size(); c++) {
    const int w = type_width(m.types[c]);
    Buffer d(ctx->arena, (size_t)std::max<int64_t>(m_out, 1) * w);
    Buffer v;
    if (m.valid[c].ptr) v = Buffer(ctx->arena, (size_t)std::max<int64_t>(m_out, 1));
    launch_gather(m.cols[c].ptr, w, rows.as<int64_t>(), m_out, d.ptr, m.valid[c].as<uint8_t>(),
                  v.ptr ? v.as<uint8_t>() : nullptr, ctx->stream);
    m.cols[c] = std::move(d);
    m.valid[c] = std::move(v);
  }
  m.rows = m_out;
}

static void topn_merge(ml_topn* T, Materialized& m);

static void topn_run(ml_topn* T, PipelineBuild& B) {
  Materialized m;
  // without operators or a selection the rows to sort are the source's own
  // columns: topn_reduce reads them in place (the caller owns them until this
  // call returns, and the winners are gathered into new buffers before that)
  bool in_place = B.P.nsteps == 0 && !B.P.sel;
  for (const VCol* v : B.layout) in_place = in_place && !v->expr && v->stage == 0;
  if (in_place) {
    for (const VCol* v : B.layout) {
      m.types.push_back(v->type);
      m.dicts.push_back(v->dict);
      Buffer d, vb;  // views: no arena, never freed
      d.ptr = const_cast<void*>(v->data);
      vb.ptr = const_cast<uint8_t*>(v->valid);
      m.cols.push_back(std::move(d));
      m.valid.push_back(std::move(vb));
    }
    m.rows = B.active_rows;
  } else {
    run_collect(B, m, true, nullptr, 0);
  }
  if (!T->schema_fixed) {
    T->best.types = m.types;
    T->best.dicts = m.dicts;
    T->best.cols.resize(m.types.size());
    T->best.valid.resize(m.types.size());
    T->best.rows = 0;
    T->schema_fixed = true;
  } else if (m.types != T->best.types) {
    fail(ML_ERR_INTERNAL, "TopN batches with different schemas");
  }
  topn_reduce(T, m);
  // nothing may keep a view of the source beyond this call (topn_reduce
  // replaced every column unless it kept no rows)
  for (size_t c = 0; c < m.cols.size(); c++) {
    if (!m.cols[c].arena) m.cols[c] = Buffer();
    if (!m.valid[c].arena) m.valid[c] = Buffer();
  }
  topn_merge(T, m);
}

// Merge reduced rows m (the same schema) into the best rows so far:
// concatenate, then reduce again.
static void topn_merge(ml_topn* T, Materialized& m) {
  Materialized& b = T->best;
  if (b.rows == 0) {
    b = std::move(m);
    return;
  }
  ml_ctx* ctx = T->ctx;
  Materialized cat;
  cat.types = b.types;
  cat.dicts = b.dicts;
  cat.rows = b.rows + m.rows;
  for (size_t c = 0; c < b.types.size(); c++) {
    const int w = type_width(b.types[c]);
    Buffer d(ctx->arena, (size_t)cat.rows * w);
    d2d(ctx, d.ptr, b.cols[c].ptr, (size_t)b.rows * w);
    d2d(ctx, d.as<uint8_t>() + b.rows * w, m.cols[c].ptr, (size_t)m.rows * w);
    Buffer v;
    if (b.valid[c].ptr || m.valid[c].ptr) {
      v = Buffer(ctx->arena, (size_t)cat.rows);
      launch_fill_u64(v.as<uint64_t>(), 0x0101010101010101ull, (cat.rows + 7) / 8, ctx->stream);
      if (b.valid[c].ptr) d2d(ctx, v.ptr, b.valid[c].ptr, (size_t)b.rows);
      if (m.valid[c].ptr) d2d(ctx, v.as<uint8_t>() + b.rows, m.valid[c].ptr, (size_t)m.rows);
    }
    cat.cols.push_back(std::move(d));
    cat.valid.push_back(std::move(v));
  }
  topn_reduce(T, cat);
  b = std::move(cat);
}

}  // namespace mlg

// ------------------------------------------------------------------ results
void ml_result::finish_views() {
  views.clear();
  for (auto& c : cols) {
    ml_column v{};
    v.type = c.type;
    v.data = c.data.ptr;
    v.len = num_rows;
    v.validity = nullptr;
    if (c.valid.ptr && num_rows > 0) {  // the exported form of the byte-per-row validity
      c.bitmap = mlg::Buffer(ctx->arena, (size_t)((num_rows + 63) / 64) * 8);
      mlg::launch_bytes_to_bits(c.valid.as<uint8_t>(), num_rows, c.bitmap.as<uint64_t>(), ctx->stream);
      v.validity = c.bitmap.as<uint64_t>();
    }
    v.dict = c.dict ? &c.dict->view : nullptr;
    views.push_back(v);
  }
}

hipEvent_t ml_ctx::pooled_event() {
  if (!event_pool.empty()) {
    hipEvent_t e = event_pool.back();
    event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  mlg::hip_check(hipEventCreate(&e), "hipEventCreate");
  return e;
}

void ml_ctx::launch(const mlg::Plan& P, int grid, int block, size_t lds) {
  if (!stats_on || P.end - P.begin < stats_min_rows) {
    if (P.nseg > 0) mlg::launch_flow(P, grid, block, lds, stream);
    else mlg::launch_pipeline(P, grid, block, lds, stream);
    return;
  }
  PendingStat ps{P.sink, P.end - P.begin, pooled_event(), pooled_event()};
  mlg::hip_check(hipEventRecord(ps.start, stream), "hipEventRecord");
  if (P.nseg > 0) mlg::launch_flow(P, grid, block, lds, stream);
  else mlg::launch_pipeline(P, grid, block, lds, stream);
  mlg::hip_check(hipEventRecord(ps.stop, stream), "hipEventRecord");
  stats.push_back(ps);
}

std::shared_ptr<mlg::HostDict> ml_ctx::dict(const ml_dict* d) {
  auto it = dicts.find(d);
  auto h = mlg::HostDict::from(d);
  if (it != dicts.end() && it->second->entries == h->entries) return it->second;
  // string equality is code equality on the device (group keys, join keys)
  std::unordered_set<std::string_view> seen;
  for (const std::string& e : h->entries)
    if (!seen.insert(e).second) mlg::fail(ML_ERR_UNSUPPORTED, "dictionary entries must be distinct");
  dicts[d] = h;
  return h;
}

// ===================================================================== C ABI
using namespace mlg;

template <class F>
static ml_status guarded(ml_ctx* ctx, F f) {
  try {
    f();
    if (ctx) ctx->last_error.clear();
    return ML_OK;
  } catch (const Error& e) {
    if (ctx) ctx->last_error = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    if (ctx) ctx->last_error = "host allocation failed";
    return ML_ERR_RESOURCES_EXHAUSTED;
  } catch (const std::exception& e) {
    if (ctx) ctx->last_error = e.what();
    return ML_ERR_INTERNAL;
  }
}

template <class Obj>
static ml_expr own_expr(Obj* o, const ml_expr& e) {
  std::vector<ml_expr_node> nodes(e.nodes, e.nodes + e.num_nodes);
  for (auto& n : nodes)
    if (n.kind == ML_EXPR_LITERAL && n.type == ML_TYPE_UTF8) o->str_store.emplace_back(n.str ? n.str : "", (size_t)n.str_len);
  o->node_store.push_back(std::move(nodes));
  auto& st = o->node_store.back();
  // string literals point at owned copies (a deque never moves its elements)
  size_t si = o->str_store.size();
  for (auto it = st.rbegin(); it != st.rend(); ++it)
    if (it->kind == ML_EXPR_LITERAL && it->type == ML_TYPE_UTF8) it->str = o->str_store[--si].c_str();
  ml_expr out = e;
  out.nodes = st.data();
  return out;
}

extern "C" {

void ml_default_config(ml_exec_config* cfg) {
  std::memset(cfg, 0, sizeof(*cfg));
  cfg->batch_rows = 1 << 21;
  cfg->device = 0;
  cfg->block_size = 256;
  cfg->grid_per_cu = 8;
  cfg->use_bloom = -1;
  cfg->stream = nullptr;
}

ml_status ml_ctx_create(const ml_exec_config* cfg, ml_ctx** out) {
  auto* ctx = new ml_ctx();
  ml_status st = guarded(nullptr, [&] {
    if (cfg) ctx->cfg = *cfg;
    else ml_default_config(&ctx->cfg);
    if (ctx->cfg.batch_rows <= 0) ctx->cfg.batch_rows = 1 << 21;
    // whole waves, up to the 256 threads the kernels are compiled for
    if (ctx->cfg.block_size != 0 && (ctx->cfg.block_size < 64 || ctx->cfg.block_size > 256 || ctx->cfg.block_size % 64))
      fail(ML_ERR_PLAN, "block_size must be 64, 128, 192 or 256");
    ctx->device = ctx->cfg.device;
    int ndev = 0;
    hip_check(hipGetDeviceCount(&ndev), "hipGetDeviceCount");
    if (ctx->device < 0 || ctx->device >= ndev) fail(ML_ERR_EXECUTION, "no such HIP device");
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    noalloc_setup(ctx->device);
    hipDeviceProp_t prop;
    hip_check(hipGetDeviceProperties(&prop, ctx->device), "hipGetDeviceProperties");
    ctx->num_cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    if (ctx->cfg.stream) {
      ctx->stream = (hipStream_t)ctx->cfg.stream;
    } else {
      hip_check(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking), "hipStreamCreate");
      ctx->own_stream = true;
    }
    ctx->arena = std::make_shared<DeviceArena>(ctx->device);
    ctx->counters = Buffer(ctx->arena, 64 * 8);
    zero(ctx, ctx->counters.ptr, 64 * 8);
    hip_check(hipHostMalloc(&ctx->pinned, ml_ctx::PINNED_BYTES, hipHostMallocDefault), "hipHostMalloc");
    ctx->checks = std::make_shared<CheckPool>();
    hip_check(hipHostMalloc((void**)&ctx->ring, ml_ctx::RING_BYTES, hipHostMallocDefault), "hipHostMalloc");
    hip_check(hipStreamSynchronize(ctx->stream), "sync");
  });
  if (st != ML_OK) {
    delete ctx;
    *out = nullptr;
    return st;
  }
  *out = ctx;
  return ML_OK;
}

void ml_ctx_destroy(ml_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  for (auto& ps : ctx->stats) {
    (void)hipEventDestroy(ps.start);
    (void)hipEventDestroy(ps.stop);
  }
  for (hipEvent_t e : ctx->event_pool) (void)hipEventDestroy(e);
  ctx->counters.reset();
  ctx->arena.reset();
  ctx->checks.reset();
  if (ctx->pinned) (void)hipHostFree(ctx->pinned);
  if (ctx->ring) (void)hipHostFree(ctx->ring);
  if (ctx->own_stream && ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

const char* ml_last_error(const ml_ctx* ctx) { return ctx ? ctx->last_error.c_str() : ""; }

ml_status ml_ctx_synchronize(ml_ctx* ctx) {
  return guarded(ctx, [&] { hip_check(hipStreamSynchronize(ctx->stream), "synchronize"); });
}

void* ml_ctx_stream(ml_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

ml_status ml_ctx_enable_stats(ml_ctx* ctx, int32_t on) {
  return guarded(ctx, [&] {
    ctx->stats_on = on != 0;
    ctx->stats_min_rows = on > 1 ? on : 0;
  });
}

int32_t ml_ctx_take_stats(ml_ctx* ctx, ml_launch_stat* out, int32_t cap) {
  int32_t n = 0;
  ml_status st = guarded(ctx, [&] {
    hip_check(hipStreamSynchronize(ctx->stream), "synchronize");
    for (auto& ps : ctx->stats) {
      float ms = 0;
      hip_check(hipEventElapsedTime(&ms, ps.start, ps.stop), "hipEventElapsedTime");
      if (n < cap) {
        out[n].sink = ps.sink;
        out[n].rows = ps.rows;
        out[n].ms = ms;
        n++;
      }
      ctx->event_pool.push_back(ps.start);
      ctx->event_pool.push_back(ps.stop);
    }
    ctx->stats.clear();
  });
  return st == ML_OK ? n : -1;
}

int64_t ml_result_num_rows(const ml_result* r) { return r ? r->num_rows : 0; }
int32_t ml_result_num_columns(const ml_result* r) { return r ? (int32_t)r->cols.size() : 0; }
const ml_column* ml_result_column(const ml_result* r, int32_t i) {
  if (!r || i < 0 || (size_t)i >= r->views.size()) return nullptr;
  return &r->views[(size_t)i];
}

ml_status ml_result_copy_column(ml_ctx* ctx, const ml_result* r, int32_t i, void* dst) {
  return guarded(ctx, [&] {
    if (!r || i < 0 || (size_t)i >= r->cols.size()) fail(ML_ERR_INTERNAL, "no such result column");
    const auto& c = r->cols[(size_t)i];
    const size_t bytes = (size_t)r->num_rows * type_width(c.type);
    if (bytes) hip_check(hipMemcpyAsync(dst, c.data.ptr, bytes, hipMemcpyDefault, ctx->stream), "copy column");
    hip_check(hipStreamSynchronize(ctx->stream), "synchronize");
  });
}

ml_status ml_result_copy_validity(ml_ctx* ctx, const ml_result* r, int32_t i, uint8_t* dst) {
  return guarded(ctx, [&] {
    if (!r || i < 0 || (size_t)i >= r->cols.size()) fail(ML_ERR_INTERNAL, "no such result column");
    const auto& c = r->cols[(size_t)i];
    if (c.valid.ptr) d2h(ctx, dst, c.valid.ptr, (size_t)r->num_rows);
    else std::memset(dst, 1, (size_t)r->num_rows);
  });
}

void ml_result_free(ml_result* r) {
  if (!r) return;
  (void)hipSetDevice(r->device);
  delete r;
}

ml_status ml_expr_type(ml_ctx* ctx, const ml_expr* expr, const ml_batch* batch, int32_t* out_type) {
  return guarded(ctx, [&] {
    PipelineBuild B(ctx);
    B.source(batch);
    *out_type = B.comp.type_of(B.resolve(*expr));
  });
}

ml_status ml_filter(ml_ctx* ctx, const ml_expr* predicate, const ml_batch* in, uint32_t* out_sel, int64_t* out_len) {
  return guarded(ctx, [&] {
    PipelineBuild B(ctx);
    B.source(in);
    ml_operator op{};
    op.kind = ML_PIPE_FILTER;
    op.predicate = *predicate;
    B.add_ops(&op, 1);
    const int64_t n = B.active_rows;
    Buffer flags(ctx->arena, (size_t)std::max<int64_t>(n, 1));
    B.P.sink = SINK_FLAGS;
    B.P.flag_out = flags.as<uint8_t>();
    B.upload();
    if (n > 0) ctx->launch(B.P, B.grid(), block_of(ctx), 0);
    const int64_t nb = compact_blocks(n);
    Buffer bc(ctx->arena, (size_t)std::max<int64_t>(nb, 1) * 4);
    Buffer offs(ctx->arena, (size_t)(nb + 1) * 8);
    launch_compact(flags.as<uint8_t>(), n, bc.as<uint32_t>(), offs.as<uint64_t>(), B.P.sel, out_sel, 1, ctx->stream);
    uint64_t total = 0;
    d2h(ctx, &total, offs.as<uint64_t>() + nb, 8);
    *out_len = (int64_t)total;
  });
}

ml_status ml_project(ml_ctx* ctx, const ml_expr* exprs, int32_t num_exprs, const ml_batch* in, ml_result** out) {
  return guarded(ctx, [&] {
    PipelineBuild B(ctx);
    B.source(in, false);  // evaluate over every physical row (eval.rs:61-66)
    if (num_exprs > MAX_COLS) fail(ML_ERR_UNSUPPORTED, "too many expressions");
    Materialized m;
    const int64_t n = in->num_rows;
    m.rows = n;
    B.P.ncols = num_exprs;
    for (int c = 0; c < num_exprs; c++) {
      NodePtr e = B.resolve(exprs[c]);
      int t = B.comp.type_of(e);
      if (t == ML_TYPE_NULL) t = ML_TYPE_INT64;
      if (t == ML_TYPE_UTF8) fail(ML_ERR_UNSUPPORTED, "string projections");
      Prog p = B.comp.compile(e, t);
      B.P.cols[c] = p;
      const int w = type_width(t);
      B.P.col_width[c] = w;
      m.types.push_back(t);
      m.dicts.push_back(nullptr);
      m.cols.emplace_back(ctx->arena, (size_t)std::max<int64_t>(n, 1) * w);
      m.valid.emplace_back();
      if (p.nullable) m.valid.back() = Buffer(ctx->arena, (size_t)std::max<int64_t>(n, 1));
      B.P.col_out[c] = m.cols.back().ptr;
      B.P.col_valid[c] = m.valid.back().ptr ? m.valid.back().as<uint8_t>() : nullptr;
    }
    B.P.sink = SINK_PROJECT;
    B.upload();
    if (n > 0) ctx->launch(B.P, B.grid(), block_of(ctx), 0);
    hip_check(hipStreamSynchronize(ctx->stream), "sync");
    *out = make_result(ctx, m);
  });
}

ml_status ml_hash_columns(ml_ctx* ctx, const ml_expr* keys, int32_t num_keys, const ml_batch* in, uint64_t* out_hashes) {
  return guarded(ctx, [&] {
    PipelineBuild B(ctx);
    B.source(in);
    if (num_keys > MAX_KEYS) fail(ML_ERR_UNSUPPORTED, "more than 4 key columns");
    B.P.nkeys = num_keys;
    B.P.key_prehashed = 0;
    for (int k = 0; k < num_keys; k++) {
      bool pre = false;
      B.P.keys[k] = B.comp.compile_hash_input(B.resolve(keys[k]), &pre);
      if (pre) B.P.key_prehashed |= 1 << k;
    }
    B.P.sink = SINK_HASH;
    B.P.hash_out = out_hashes;
    B.upload();
    if (B.active_rows > 0) ctx->launch(B.P, B.grid(), block_of(ctx), 0);
    hip_check(hipStreamSynchronize(ctx->stream), "sync");
  });
}

// ------------------------------------------------------------- hash aggregate
ml_status ml_hashagg_create(ml_ctx* ctx, const ml_agg_spec* spec, ml_hashagg** out) {
  auto* A = new ml_hashagg();
  A->ctx = ctx;
  A->pool = ctx->mem;
  ml_status st = guarded(ctx, [&] {
    if (spec->num_keys < 0 || spec->num_aggs < 0) fail(ML_ERR_PLAN, "bad aggregate spec");
    if (spec->mode < ML_AGG_MODE_SINGLE || spec->mode > ML_AGG_MODE_FINAL) fail(ML_ERR_PLAN, "bad aggregate mode");
    A->mode = spec->mode;
    for (int a = 0; a < spec->num_aggs; a++) {
      AggDef u;
      u.func = spec->aggs[a].func;
      u.arg_type = spec->aggs[a].arg_type;
      u.raw = nullptr;
      A->user_aggs.push_back(u);
    }
    // Final mode reads its input by position, group keys first, then each
    // aggregate's state columns (hash.rs:125-137, ungrouped.rs:100-113): the
    // group expressions and aggregate arguments are replaced by column refs.
    auto column_expr = [&](int c) {
      ml_expr_node n{};
      n.kind = ML_EXPR_COLUMN;
      n.column = c;
      ml_expr e{};
      e.nodes = &n;
      e.num_nodes = 1;
      return own_expr(A, e);
    };
    const bool final_mode = A->mode == ML_AGG_MODE_FINAL;
    for (int k = 0; k < spec->num_keys; k++) {
      A->group_raw.push_back(final_mode ? column_expr(k) : own_expr(A, spec->group_exprs[k]));
      parse_expr(A->group_raw.back());
      A->key_types.push_back(spec->key_types[k]);
    }
    if (final_mode) {
      int col = spec->num_keys;
      for (const AggDef& u : A->user_aggs) {
        A->state_col.push_back((int)A->aggs.size());
        for (int t : agg_state_types(u.func, u.arg_type)) {
          AggDef d;
          d.func = ML_AGG_SUM;
          d.arg_type = t;
          A->group_raw.push_back(column_expr(col++));
          d.arg = parse_expr(A->group_raw.back());
          d.raw = nullptr;
          A->aggs.push_back(d);
        }
      }
    } else {
      for (int a = 0; a < spec->num_aggs; a++) {
        AggDef d = A->user_aggs[(size_t)a];
        if (d.func != ML_AGG_COUNT_STAR) {
          A->group_raw.push_back(own_expr(A, spec->aggs[a].arg));
          d.arg = parse_expr(A->group_raw.back());
        }
        A->aggs.push_back(d);
      }
    }
    // point each argument source at its stored expression (after all push_backs)
    size_t gi = (size_t)spec->num_keys;
    for (auto& d : A->aggs)
      if (final_mode || d.func != ML_AGG_COUNT_STAR) d.raw = &A->group_raw[gi++];
    A->sizing = spec->sizing;
    A->hint = spec->capacity_hint;
    if (A->sizing != ML_SIZING_PLAIN && A->hint <= 0) A->hint = 1;
  });
  if (st != ML_OK) {
    delete A;
    *out = nullptr;
    return st;
  }
  *out = A;
  return ML_OK;
}

ml_status ml_hashagg_sink(ml_hashagg* agg, const ml_batch* batch) {
  return guarded(agg->ctx, [&] {
    PipelineBuild B(agg->ctx);
    B.source(batch);
    agg_run(agg, B);
  });
}

ml_status ml_pipeline_hashagg(ml_ctx* ctx, const ml_batch* source, const ml_operator* ops, int32_t num_ops,
                              ml_hashagg* sink) {
  return guarded(ctx, [&] {
    const bool fresh = !sink->planned;
    {
      PipelineBuild B(ctx);
      B.source(source);
      B.add_ops(ops, num_ops);
      agg_run(sink, B);
    }
    if (!sink->direct && sink->sizing != ML_SIZING_PLAIN) {
      // a pre-sized table that filled up lost rows: the learned path reruns
      // this batch with the plain resize-on-overflow sizing
      if (agg_overflow_bits(sink) & 2u) {
        if (sink->sizing == ML_SIZING_HINT) fail(ML_ERR_CAPACITY, "pre-sized group table overflowed");
        if (!fresh) fail(ML_ERR_CAPACITY, "learned-size table overflowed after earlier batches");
        agg_reset_table(sink);
        sink->sizing = ML_SIZING_PLAIN;
        sink->resizes = 0;
        PipelineBuild B(ctx);
        B.source(source);
        B.add_ops(ops, num_ops);
        agg_run(sink, B);
      }
    }
  });
}

ml_status ml_hashagg_finalize(ml_hashagg* agg, ml_result** out) {
  return guarded(agg->ctx, [&] {
    if (agg->merger) {  // combined per-thread states (and this sink's own groups)
      std::lock_guard<std::mutex> lock(agg->combine_mu);
      std::unique_ptr<ml_result> own(agg_states(agg));
      if (own) agg_merge_states(agg, own.get());
      ml_hashagg* M = agg->merger;
      *out = M->spilled ? agg_merge_spilled(M) : agg_finalize(M, false);
      return;
    }
    *out = agg->spilled ? agg_merge_spilled(agg) : agg_finalize(agg, false);
  });
}

ml_status ml_ctx_set_memory_limit(ml_ctx* ctx, int64_t bytes, int32_t spill) {
  return guarded(ctx, [&] {
    if (bytes < 0) fail(ML_ERR_PLAN, "negative memory limit");
    ctx->mem->limit = bytes;
    ctx->mem->spill = spill != 0;
    ctx->mem->peak = ctx->mem->used;  // the peak is counted from the new limit on
  });
}

void ml_ctx_memory_usage(const ml_ctx* ctx, int64_t* used, int64_t* peak, int64_t* limit) {
  if (used) *used = ctx->mem->used;
  if (peak) *peak = ctx->mem->peak;
  if (limit) *limit = ctx->mem->limit;
}

ml_status ml_hashagg_num_groups(ml_hashagg* agg, int64_t* out) {
  return guarded(agg->ctx, [&] {
    if (agg->direct) {
      const size_t ns = agg->slots.size();
      std::vector<uint64_t> tot((size_t)agg->ngroups * ns);
      d2h(agg->ctx, tot.data(), agg->totals.ptr, tot.size() * 8);
      int64_t n = 0;
      for (int64_t g = 0; g < agg->ngroups; g++) n += tot[(size_t)g * ns] > 0 || agg->key_types.empty();
      *out = n;
    } else {
      agg->len_known = false;
      *out = agg->planned ? (int64_t)agg_len(agg) : 0;
    }
  });
}

ml_status ml_hashagg_table_stats(ml_hashagg* agg, int64_t* capacity, int32_t* resizes) {
  return guarded(agg->ctx, [&] {
    *capacity = agg->direct ? agg->ngroups : (int64_t)agg->cap;
    *resizes = agg->resizes;
  });
}

void ml_hashagg_free(ml_hashagg* agg) { delete agg; }

// ------------------------------------------------------------------ join
ml_status ml_join_build_create(ml_ctx* ctx, const ml_join_build_spec* spec, ml_join_build** out) {
  auto* J = new ml_join_build();
  J->ctx = ctx;
  ml_status st = guarded(ctx, [&] {
    for (int k = 0; k < spec->num_keys; k++) {
      J->keys_raw.push_back(own_expr(J, spec->keys[k]));
      parse_expr(J->keys_raw.back());
      J->key_types.push_back(spec->key_types[k]);
      if (spec->key_types[k] < ML_TYPE_BOOLEAN || spec->key_types[k] > ML_TYPE_UTF8) fail(ML_ERR_PLAN, "bad key type");
    }
  });
  if (st != ML_OK) {
    delete J;
    *out = nullptr;
    return st;
  }
  *out = J;
  return ML_OK;
}

ml_status ml_join_build_sink(ml_join_build* b, const ml_batch* batch) {
  return guarded(b->ctx, [&] {
    PipelineBuild B(b->ctx);
    B.source(batch);
    join_build_run(b, B);
  });
}

ml_status ml_pipeline_join_build(ml_ctx* ctx, const ml_batch* source, const ml_operator* ops, int32_t num_ops,
                                 ml_join_build* sink) {
  return guarded(ctx, [&] {
    PipelineBuild B(ctx);
    B.source(source);
    B.add_ops(ops, num_ops);
    join_build_run(sink, B);
  });
}

// The table (rows, key structures, filters) held against the budget
// (JoinBuildSink::finalize, join.rs:120-153).
static void reserve_join_table(ml_ctx* ctx, ml_join_table* t) {
  const JoinTableImpl& J = t->impl;
  int64_t bytes = (int64_t)(J.slots.bytes + J.bloom.bytes + J.dense_bits.bytes + J.dense_rank.bytes +
                            J.dense_keys.bytes + J.keys.bytes + J.key_null.bytes);
  for (auto* v : {&J.cols, &J.col_valid, &J.orig_cols, &J.orig_valid})
    for (const Buffer& x : *v) bytes += (int64_t)x.bytes;
  if (!ctx->mem->try_grow(bytes))
    fail(ML_ERR_RESOURCES_EXHAUSTED, "HashJoinBuild(table): " + std::to_string(bytes) +
                                         " bytes do not fit the memory budget (" + std::to_string(ctx->mem->used) +
                                         " of " + std::to_string(ctx->mem->limit) + " bytes reserved)");
  t->pool = ctx->mem;
  t->reserved = bytes;
}

ml_status ml_join_build_finalize(ml_join_build* b, ml_join_table** out) {
  ml_ctx* ctx = b->ctx;
  ml_status st = guarded(ctx, [&] {
    if (!b->schema_fixed) fail(ML_ERR_EXECUTION, "join build received no batch");
    FinalizePrep P;
    join_finalize_enqueue(b, P, (int64_t*)ctx->pinned);
    if (P.small || P.ranged) {
      const uint64_t seq = ctx->check_seq;
      hip_check(hipStreamSynchronize(ctx->stream), "synchronize");
      ctx->synced_seq = seq;
      std::memcpy(P.slot, ctx->pinned, 32);
    }
    std::unique_ptr<ml_join_table> t(join_finalize_finish(b, P));
    reserve_join_table(ctx, t.get());
    *out = t.release();
  });
  delete b;
  return st;
}

ml_status ml_join_build_finalize_many(ml_join_build** builds, int32_t n, ml_join_table** out) {
  if (n <= 0) return ML_OK;
  ml_ctx* ctx = builds[0]->ctx;
  std::vector<std::unique_ptr<ml_join_table>> tables((size_t)n);
  ml_status st = guarded(ctx, [&] {
    if (n > (int32_t)(ml_ctx::PINNED_BYTES / 32)) fail(ML_ERR_PLAN, "too many join builds in one call");
    std::vector<FinalizePrep> P((size_t)n);
    for (int32_t i = 0; i < n; i++) {
      if (builds[i]->ctx != ctx) fail(ML_ERR_PLAN, "join builds of one call must share a context");
      if (!builds[i]->schema_fixed) fail(ML_ERR_EXECUTION, "join build received no batch");
      P[(size_t)i].batched = n > 1;
      join_finalize_enqueue(builds[i], P[(size_t)i], (int64_t*)ctx->pinned + 4 * i);
    }
    const uint64_t seq = ctx->check_seq;
    hip_check(hipStreamSynchronize(ctx->stream), "synchronize");
    ctx->synced_seq = seq;
    for (int32_t i = 0; i < n; i++) std::memcpy(P[(size_t)i].slot, (int64_t*)ctx->pinned + 4 * i, 32);
    for (int32_t i = 0; i < n; i++) {
      tables[(size_t)i].reset(join_finalize_finish(builds[i], P[(size_t)i]));
      reserve_join_table(ctx, tables[(size_t)i].get());
    }
  });
  for (int32_t i = 0; i < n; i++) {
    out[i] = st == ML_OK ? tables[(size_t)i].release() : nullptr;
    delete builds[i];
  }
  return st;
}

int64_t ml_join_table_num_rows(const ml_join_table* t) { return t ? t->impl.nrows : 0; }
// Settle a pending build check for an accessor without a status code.
static bool resolved(const ml_join_table* t) {
  if (!t) return false;
  return guarded(t->ctx, [&] { join_table_resolve(const_cast<ml_join_table*>(t)); }) == ML_OK;
}

namespace mlg {
// The rank of an offset-indexed table's presence bits (row numbers in key order).
static const uint32_t* offset_rank(ml_ctx* ctx, JoinTableImpl& T) {
  if (!T.offset_rank.ptr) {
    const int64_t ne = (int64_t)((T.dev.dense_range + 63) / 64);
    T.offset_rank = Buffer(ctx->arena, (size_t)(ne + 1) * 4);
    Buffer sums(ctx->arena, (size_t)(scan_u32_blocks(ne) + 1) * 4);
    launch_dense_rank(T.dense_bits.as<uint64_t>(), ne, T.offset_rank.as<uint32_t>(), sums.as<uint32_t>(), ctx->stream);
  }
  return T.offset_rank.as<uint32_t>();
}

// Columns of an offset-indexed table in key order (the ABI's row numbering), made once.
// The first-key column of a table that does not store it, rebuilt at its
// row positions (p < range: min + p; NULL-key rows after: 0) for the views.
static void materialize_implied(ml_ctx* ctx, JoinTableImpl& T) {
  if (T.implied_col < 0 || T.cols[(size_t)T.implied_col].ptr) return;
  const size_t c = (size_t)T.implied_col;
  const int w = type_width(T.col_types[c]);
  const int64_t held = T.dense_offset ? T.offset_rows : T.nrows;  // a full table holds nrows = range + NULL rows
  T.cols[c] = Buffer(ctx->arena, (size_t)std::max<int64_t>(held, 1) * w);
  launch_dense_key_fill(T.cols[c].ptr, w, held, T.dev.dense_range, T.dev.dense_min, ctx->stream);
  refresh_views(T);
}

static void compact_offset_table(ml_ctx* ctx, JoinTableImpl& T) {
  if (!T.dense_offset || !T.compact_cols.empty()) return;
  const uint32_t* rank = offset_rank(ctx, T);
  for (size_t c = 0; c < T.cols.size(); c++) {
    const int w = type_width(T.col_types[c]);
    T.compact_cols.emplace_back(ctx->arena, (size_t)std::max<int64_t>(T.nrows, 1) * w);
    T.compact_valid.emplace_back();
    if (T.col_valid[c].ptr) T.compact_valid.back() = Buffer(ctx->arena, (size_t)std::max<int64_t>(T.nrows, 1));
    launch_dense_compact(T.cols[c].ptr, w, T.col_valid[c].ptr ? T.col_valid[c].as<uint8_t>() : nullptr,
                         T.offset_rows, T.dense_bits.as<uint64_t>(), rank, T.dev.dense_range,
                         T.compact_cols.back().ptr,
                         T.compact_valid.back().ptr ? T.compact_valid.back().as<uint8_t>() : nullptr, ctx->stream);
  }
  refresh_views(T);
}
}  // namespace mlg
int32_t ml_join_table_unique(const ml_join_table* t) { return resolved(t) ? t->impl.dev.unique : 0; }
int32_t ml_join_table_num_columns(const ml_join_table* t) { return t ? (int32_t)t->impl.col_types.size() : 0; }
const ml_column* ml_join_table_column(const ml_join_table* t, int32_t i) {
  if (!resolved(t) || i < 0 || (size_t)i >= t->impl.views.size()) return nullptr;
  if (guarded(t->ctx, [&] {
        JoinTableImpl& T = const_cast<ml_join_table*>(t)->impl;
        materialize_implied(t->ctx, T);
        compact_offset_table(t->ctx, T);
      }) != ML_OK)
    return nullptr;
  // the build's last kernels may still be queued: the columns are complete
  // when this returns, so they can be read from any stream
  if (hipStreamSynchronize(t->ctx->stream) != hipSuccess) return nullptr;
  return &t->impl.views[(size_t)i];
}
void ml_join_table_free(ml_join_table* t) { delete t; }

ml_status ml_join_probe(ml_ctx* ctx, const ml_join_table* t, const ml_expr* probe_keys, int32_t num_keys,
                        const ml_batch* probe, ml_result** out_pairs) {
  return guarded(ctx, [&] {
    PipelineBuild B(ctx);
    B.source(probe);
    ml_operator op{};
    op.kind = ML_PIPE_PROBE;
    op.table = t;
    op.keys = probe_keys;
    op.num_keys = num_keys;
    B.add_ops(&op, 1);
    Materialized m;
    Buffer pairs;
    run_collect(B, m, false, &pairs, 2);
    // an offset-indexed table's rows are key offsets here: renumber them in key order
    JoinTableImpl& J = const_cast<ml_join_table*>(t)->impl;
    if (J.dense_offset && m.rows > 0)
      launch_dense_translate(pairs.as<int64_t>() + 1, 2, (int64_t)m.rows, J.dense_bits.as<uint64_t>(),
                             offset_rank(ctx, J), J.dev.dense_range, ctx->stream);
    // split the (probe row, build row) pairs into two INT64 columns
    Materialized o;
    o.rows = m.rows;
    for (int c = 0; c < 2; c++) {
      o.types.push_back(ML_TYPE_INT64);
      o.dicts.push_back(nullptr);
      o.cols.emplace_back(ctx->arena, (size_t)std::max<int64_t>(m.rows, 1) * 8);
      o.valid.emplace_back();
    }
    if (m.rows > 0) {
      hip_check(hipMemcpy2DAsync(o.cols[0].ptr, 8, pairs.ptr, 16, 8, (size_t)m.rows, hipMemcpyDeviceToDevice, ctx->stream), "split");
      hip_check(hipMemcpy2DAsync(o.cols[1].ptr, 8, pairs.as<uint8_t>() + 8, 16, 8, (size_t)m.rows, hipMemcpyDeviceToDevice,
                                 ctx->stream), "split");
    }
    hip_check(hipStreamSynchronize(ctx->stream), "sync");
    *out_pairs = make_result(ctx, o);
  });
}

ml_status ml_pipeline_collect(ml_ctx* ctx, const ml_batch* source, const ml_operator* ops, int32_t num_ops,
                              ml_result** out) {
  return guarded(ctx, [&] {
    PipelineBuild B(ctx);
    B.source(source);
    B.add_ops(ops, num_ops);
    Materialized m;
    run_collect(B, m, true, nullptr, 0);
    *out = make_result(ctx, m);
  });
}

ml_status ml_pipeline_limit(ml_ctx* ctx, const ml_batch* source, const ml_operator* ops, int32_t num_ops,
                            int64_t limit, int64_t offset, ml_result** out) {
  return guarded(ctx, [&] {
    if (limit < 0 || offset < 0) fail(ML_ERR_PLAN, "negative LIMIT or OFFSET");
    PipelineBuild B(ctx);
    B.source(source);
    B.add_ops(ops, num_ops);
    Materialized m;
    run_collect(B, m, true, nullptr, 0);
    // limit.rs:64-83: keep rows [offset, offset + limit)
    const int64_t start = std::min(offset, m.rows);
    const int64_t len = std::min(limit, m.rows - start);
    if (start > 0 || len < m.rows) {
      for (size_t c = 0; c < m.cols.size(); c++) {
        const int w = type_width(m.types[c]);
        Buffer d(ctx->arena, (size_t)std::max<int64_t>(len, 1) * w);
        d2d(ctx, d.ptr, m.cols[c].as<uint8_t>() + start * w, (size_t)len * w);
        m.cols[c] = std::move(d);
        if (m.valid[c].ptr) {
          Buffer v(ctx->arena, (size_t)std::max<int64_t>(len, 1));
          d2d(ctx, v.ptr, m.valid[c].as<uint8_t>() + start, (size_t)len);
          m.valid[c] = std::move(v);
        }
      }
      m.rows = len;
    }
    *out = make_result(ctx, m);
  });
}

// ------------------------------------------------------------- MLB1 bytes
// write_batch (ipc.rs:44-87): per column a 2-byte header (type tag = the
// ml_type value, validity flag), the validity words when a kept row is NULL,
// then the payload; strings as u32 offsets, u64 byte count and the bytes.
ml_status ml_mlb1_write(ml_ctx* ctx, const ml_batch* b, void* dst, int64_t cap, int64_t* size) {
  return guarded(ctx, [&] {
    if (!b || !size) fail(ML_ERR_INTERNAL, "null argument");
    hipStream_t s = ctx->stream;
    const uint32_t* sel = b->selection;
    const int64_t n = sel ? b->selection_len : b->num_rows;
    const int nc = b->num_columns;
    struct Prep {
      Buffer valid, offs, dict_off, dict_bytes;
      uint32_t ndict = 0;
    };
    std::vector<Prep> prep((size_t)std::max(nc, 0));
    Buffer stats(ctx->arena, (size_t)std::max(nc, 1) * 16);  // per column: NULL rows, string bytes
    launch_fill_u64(stats.as<uint64_t>(), 0, (int64_t)std::max(nc, 1) * 2, s);
    for (int c = 0; c < nc; c++) {
      const ml_column& mc = b->columns[c];
      Prep& P = prep[(size_t)c];
      if (mc.type < ML_TYPE_BOOLEAN || mc.type > ML_TYPE_UTF8) fail(ML_ERR_PLAN, "bad column type");
      if (mc.len != b->num_rows) fail(ML_ERR_INTERNAL, "column length differs from batch rows");
      if (n > 0 && !mc.data) fail(ML_ERR_INTERNAL, "null column data");
      if (mc.validity && n > 0) {
        P.valid = Buffer(ctx->arena, (size_t)b->num_rows);
        launch_bits_to_bytes(mc.validity, b->num_rows, P.valid.as<uint8_t>(), s);
        launch_null_count(P.valid.as<uint8_t>(), sel, n, stats.as<uint64_t>() + 2 * c, s);
      }
      if (mc.type == ML_TYPE_UTF8) {
        if (!mc.dict) fail(ML_ERR_UNSUPPORTED, "UTF8 columns must be dictionary-encoded");
        auto d = HostDict::from(mc.dict);
        P.ndict = (uint32_t)d->entries.size();
        P.dict_off = Buffer(ctx->arena, d->offsets.size() * 4);
        upload(ctx, P.dict_off.ptr, d->offsets.data(), d->offsets.size() * 4);
        P.dict_bytes = Buffer(ctx->arena, std::max<size_t>(d->bytes.size(), 1));
        upload(ctx, P.dict_bytes.ptr, d->bytes.data(), d->bytes.size());
        Buffer len(ctx->arena, (size_t)std::max<int64_t>(n, 1) * 4);
        launch_mlb1_str_len((const uint32_t*)mc.data, sel, n, P.dict_off.as<uint32_t>(), P.ndict,
                            len.as<uint32_t>(), s);
        P.offs = Buffer(ctx->arena, (size_t)(n + 1) * 4);
        if (n > 0) {
          Buffer sums(ctx->arena, (size_t)(scan_u32_blocks(n) + 1) * 4);
          launch_scan_u32(len.as<uint32_t>(), n, P.offs.as<uint32_t>(), sums.as<uint32_t>(), s);
          d2d(ctx, stats.as<uint64_t>() + 2 * c + 1, P.offs.as<uint32_t>() + n, 4);  // low half (little-endian)
        } else {
          launch_fill_u64(stats.as<uint64_t>() + 2 * c + 1, 0, 1, s);
          const uint32_t zero = 0;
          upload(ctx, P.offs.ptr, &zero, 4);
        }
      }
    }
    std::vector<uint64_t> st((size_t)std::max(nc, 1) * 2);
    d2h(ctx, st.data(), stats.ptr, st.size() * 8);
    // layout
    const int64_t words = (n + 63) / 64;
    int64_t total = 16;
    for (int c = 0; c < nc; c++) {
      const int t = b->columns[c].type;
      total += 2 + (st[2 * c] ? words * 8 : 0);
      total += t == ML_TYPE_UTF8 ? (n + 1) * 4 + 8 + (int64_t)st[2 * c + 1] : n * type_width(t);
    }
    *size = total;
    if (!dst) return;
    if (cap < total) fail(ML_ERR_CAPACITY, "MLB1 destination smaller than the batch");
    uint8_t* out = static_cast<uint8_t*>(dst);
    uint8_t hdr[16] = {'M', 'L', 'B', '1'};
    const uint32_t nc32 = (uint32_t)nc;
    const uint64_t n64 = (uint64_t)n;
    std::memcpy(hdr + 4, &nc32, 4);
    std::memcpy(hdr + 8, &n64, 8);
    upload(ctx, out, hdr, 16);
    int64_t pos = 16;
    for (int c = 0; c < nc; c++) {
      const ml_column& mc = b->columns[c];
      const Prep& P = prep[(size_t)c];
      const uint8_t ch[2] = {(uint8_t)mc.type, (uint8_t)(st[2 * c] ? 1 : 0)};
      upload(ctx, out + pos, ch, 2);
      pos += 2;
      if (st[2 * c]) {
        launch_mlb1_pack_validity(P.valid.as<uint8_t>(), sel, n, out + pos, s);
        pos += words * 8;
      }
      if (mc.type == ML_TYPE_UTF8) {
        launch_mlb1_pack_fixed(P.offs.ptr, 4, false, nullptr, n + 1, out + pos, s);
        pos += (n + 1) * 4;
        upload(ctx, out + pos, &st[2 * c + 1], 8);
        pos += 8;
        launch_mlb1_pack_strings((const uint32_t*)mc.data, sel, n, P.dict_off.as<uint32_t>(),
                                 P.dict_bytes.as<uint8_t>(), P.ndict, P.offs.as<uint32_t>(), out + pos, s);
        pos += (int64_t)st[2 * c + 1];
      } else {
        const int w = type_width(mc.type);
        launch_mlb1_pack_fixed(mc.data, w, mc.type == ML_TYPE_BOOLEAN, sel, n, out + pos, s);
        pos += n * w;
      }
    }
    if (pos != total) fail(ML_ERR_INTERNAL, "MLB1 layout mismatch");
    hip_check(hipStreamSynchronize(s), "synchronize");
  });
}

// read_batch (ipc.rs:111-170) from host bytes into device columns; strings
// are dictionary-encoded again (entries in first-seen order).
ml_status ml_mlb1_read(ml_ctx* ctx, const void* src, int64_t len, int64_t* consumed, ml_result** out) {
  if (out) *out = nullptr;
  if (consumed) *consumed = 0;
  return guarded(ctx, [&] {
    if (!out || !consumed || (len > 0 && !src)) fail(ML_ERR_INTERNAL, "null argument");
    const uint8_t* p = static_cast<const uint8_t*>(src);
    const size_t avail = (size_t)std::max<int64_t>(len, 0);
    size_t pos = 0;
    auto take = [&](size_t k) {
      if (avail - pos < k) fail(ML_ERR_IO, "failed to fill whole buffer");
      const uint8_t* q = p + pos;
      pos += k;
      return q;
    };
    auto u32 = [&]() { uint32_t x; std::memcpy(&x, take(4), 4); return x; };
    auto u64 = [&]() { uint64_t x; std::memcpy(&x, take(8), 8); return x; };
    if (avail < 4) return;  // a clean end of stream
    if (std::memcmp(take(4), "MLB1", 4) != 0) fail(ML_ERR_INTERNAL, "bad batch magic");
    const uint32_t nc = u32();
    const uint64_t n64 = u64();
    if (n64 > (uint64_t)INT64_MAX / 8) fail(ML_ERR_IO, "failed to fill whole buffer");
    const int64_t n = (int64_t)n64;
    std::unique_ptr<ml_result> r(new ml_result());
    r->ctx = ctx;
    r->device = ctx->device;
    r->num_rows = n;
    for (uint32_t ci = 0; ci < nc; ci++) {
      const uint8_t* hdr = take(2);
      if (hdr[0] > 5) fail(ML_ERR_INTERNAL, "bad type tag " + std::to_string(hdr[0]));
      ResultColumn c;
      c.type = hdr[0];
      if (hdr[1] == 1) {
        const size_t words = (size_t)(n + 63) / 64;
        const uint8_t* q = take(words * 8);
        if (n > 0) {
          Buffer wb(ctx->arena, words * 8);
          upload(ctx, wb.ptr, q, words * 8);
          c.valid = Buffer(ctx->arena, (size_t)n);
          launch_bits_to_bytes(wb.as<uint64_t>(), n, c.valid.as<uint8_t>(), ctx->stream);
        }
      }
      if (c.type == ML_TYPE_UTF8) {
        const uint8_t* oq = take(((size_t)n + 1) * 4);
        std::vector<uint32_t> offs((size_t)n + 1);
        std::memcpy(offs.data(), oq, offs.size() * 4);
        const uint64_t blen = u64();
        const uint8_t* bytes = take((size_t)blen);
        bool ok = offs[0] == 0 && offs[(size_t)n] == blen;
        for (int64_t k = 0; k < n && ok; k++) ok = offs[(size_t)k] <= offs[(size_t)k + 1];
        if (!ok) fail(ML_ERR_INTERNAL, "corrupt string offsets");
        auto d = std::make_shared<HostDict>();
        std::unordered_map<std::string_view, uint32_t> code;
        std::vector<uint32_t> codes((size_t)n);
        for (int64_t k = 0; k < n; k++) {
          const std::string_view v(reinterpret_cast<const char*>(bytes) + offs[(size_t)k],
                                   offs[(size_t)k + 1] - offs[(size_t)k]);
          auto it = code.find(v);
          if (it == code.end()) {
            it = code.emplace(v, (uint32_t)d->entries.size()).first;
            d->entries.emplace_back(v);
          }
          codes[(size_t)k] = it->second;
        }
        d->finish_view();
        c.dict = d;
        c.data = Buffer(ctx->arena, (size_t)std::max<int64_t>(n, 1) * 4);
        upload(ctx, c.data.ptr, codes.data(), codes.size() * 4);
      } else {
        const int w = type_width(c.type);
        const uint8_t* q = take((size_t)n * w);
        c.data = Buffer(ctx->arena, (size_t)std::max<int64_t>(n, 1) * w);
        if (c.type == ML_TYPE_BOOLEAN) {
          std::vector<uint8_t> v(q, q + n);
          for (auto& x : v) x = x != 0;
          upload(ctx, c.data.ptr, v.data(), v.size());
        } else {
          upload(ctx, c.data.ptr, q, (size_t)n * w);
        }
      }
      r->cols.push_back(std::move(c));
    }
    r->finish_views();
    hip_check(hipStreamSynchronize(ctx->stream), "synchronize");
    *consumed = (int64_t)pos;
    *out = r.release();
  });
}

ml_status ml_pipeline_collect_sized(ml_ctx* ctx, const ml_batch* source, const ml_operator* ops, int32_t num_ops,
                                    int64_t capacity_hint, int32_t* used_fallback, ml_result** out) {
  return guarded(ctx, [&] {
    if (used_fallback) *used_fallback = 0;
    {
      PipelineBuild B(ctx);
      B.source(source);
      B.add_ops(ops, num_ops);
      Materialized m;
      if (run_collect(B, m, true, nullptr, 0, std::max<int64_t>(capacity_hint, 0))) {
        *out = make_result(ctx, m);
        return;
      }
    }
    if (used_fallback) *used_fallback = 1;
    PipelineBuild B(ctx);
    B.source(source);
    B.add_ops(ops, num_ops);
    Materialized m;
    run_collect(B, m, true, nullptr, 0);
    *out = make_result(ctx, m);
  });
}

ml_status ml_pipeline_count(ml_ctx* ctx, const ml_batch* source, const ml_operator* ops, int32_t num_ops,
                            int64_t* out_rows) {
  return guarded(ctx, [&] {
    PipelineBuild B(ctx);
    B.source(source);
    B.add_ops(ops, num_ops);
    *out_rows = (int64_t)run_count(B);
  });
}

// -------------------------------------------------------------------- TopN
ml_status ml_topn_create(ml_ctx* ctx, const ml_sort_spec* spec, ml_topn** out) {
  auto* T = new ml_topn();
  T->ctx = ctx;
  ml_status st = guarded(ctx, [&] {
    for (int k = 0; k < spec->num_keys; k++) {
      T->keys_raw.push_back(own_expr(T, spec->keys[k].expr));
      parse_expr(T->keys_raw.back());
      T->asc.push_back(spec->keys[k].asc != 0);
      T->nulls_first.push_back(spec->keys[k].nulls_first != 0);
    }
    T->limit = spec->limit < 0 ? -1 : spec->limit;  // -1: ORDER BY without LIMIT
  });
  if (st != ML_OK) {
    delete T;
    *out = nullptr;
    return st;
  }
  *out = T;
  return ML_OK;
}

ml_status ml_topn_sink(ml_topn* t, const ml_batch* batch) {
  return guarded(t->ctx, [&] {
    PipelineBuild B(t->ctx);
    B.source(batch);
    topn_run(t, B);
  });
}

ml_status ml_pipeline_topn(ml_ctx* ctx, const ml_batch* source, const ml_operator* ops, int32_t num_ops, ml_topn* sink) {
  return guarded(ctx, [&] {
    PipelineBuild B(ctx);
    B.source(source);
    B.add_ops(ops, num_ops);
    topn_run(sink, B);
  });
}

ml_status ml_topn_finalize(ml_topn* t, ml_result** out) {
  return guarded(t->ctx, [&] {
    Materialized& b = t->best;
    if (!t->schema_fixed) b.rows = 0;
    *out = make_result(t->ctx, b);
    t->best = Materialized();
    t->schema_fixed = false;
  });
}

void ml_topn_free(ml_topn* t) { delete t; }

// ------------------------------------------------------ per-thread sinks
// LocalSink (crates/minilake-exec/src/pipeline.rs:70-79): create_local makes
// a sink state on the calling thread's context; combine merges it into the
// global sink under the global's lock and frees it.

}  // extern "C"

namespace mlg {

// The specification a hash aggregate was created with (its copied expressions).
static void agg_spec_of(ml_hashagg* A, std::vector<ml_aggregate>& aggs, ml_agg_spec& spec) {
  const size_t nk = A->key_types.size();
  aggs.assign(A->user_aggs.size(), ml_aggregate{});
  for (size_t a = 0; a < aggs.size(); a++) {
    aggs[a].func = A->user_aggs[a].func;
    aggs[a].arg_type = A->user_aggs[a].arg_type;
    if (A->mode != ML_AGG_MODE_FINAL && aggs[a].func != ML_AGG_COUNT_STAR) aggs[a].arg = *A->aggs[a].raw;
  }
  spec = ml_agg_spec{};
  spec.group_exprs = A->group_raw.data();
  spec.key_types = A->key_types.data();
  spec.num_keys = (int32_t)nk;
  spec.num_aggs = (int32_t)aggs.size();
  spec.aggs = aggs.data();
  spec.mode = A->mode;
  spec.sizing = A->sizing;
  spec.capacity_hint = A->sizing == ML_SIZING_PLAIN ? 0 : A->hint;
}

// The partial states of an aggregate's groups (Accumulator::state), also
// after a spill; null when it never received a batch.
static ml_result* agg_states(ml_hashagg* A) {
  if (!A->planned) return nullptr;
  if (A->spilled) {
    A->states_out = true;
    return agg_merge_spilled(A);
  }
  return agg_finalize(A, true);
}

// Sink a states result into the global's merger (a Final aggregate of the
// same groups and aggregates, created at the first merge).
static void agg_merge_states(ml_hashagg* G, ml_result* st) {
  ml_ctx* ctx = G->ctx;
  if (!G->merger) {
    std::vector<ml_aggregate> aggs;
    ml_agg_spec spec;
    agg_spec_of(G, aggs, spec);
    spec.mode = ML_AGG_MODE_FINAL;
    ml_hashagg* M = nullptr;
    const ml_status rc = ml_hashagg_create(ctx, &spec, &M);
    if (rc != ML_OK) fail(rc, ctx->last_error);
    M->emit_states = G->mode == ML_AGG_MODE_PARTIAL;
    G->merger = M;
  }
  ml_batch b{};
  b.columns = st->views.data();
  b.num_columns = (int32_t)st->views.size();
  b.num_rows = st->num_rows;
  const ml_status rc = ml_hashagg_sink(G->merger, &b);
  if (rc != ML_OK) fail(rc, ctx->last_error);
}

}  // namespace mlg

// The result of a slice [start, start + len) of r's rows.
static ml_result* slice_result(ml_ctx* ctx, ml_result* r, int64_t start, int64_t len) {
  auto* o = new ml_result();
  o->ctx = ctx;
  o->device = ctx->device;
  o->num_rows = len;
  for (auto& c : r->cols) {
    ResultColumn rc;
    rc.type = c.type;
    rc.dict = c.dict;
    const int w = type_width(c.type);
    rc.data = Buffer(ctx->arena, (size_t)std::max<int64_t>(len, 1) * w);
    if (len > 0) d2d(ctx, rc.data.ptr, c.data.as<uint8_t>() + start * w, (size_t)len * w);
    if (c.valid.ptr) {
      rc.valid = Buffer(ctx->arena, (size_t)std::max<int64_t>(len, 1));
      if (len > 0) d2d(ctx, rc.valid.ptr, c.valid.as<uint8_t>() + start, (size_t)len);
    }
    o->cols.push_back(std::move(rc));
  }
  o->finish_views();
  hip_check(hipStreamSynchronize(ctx->stream), "synchronize");
  return o;
}

// LimitSink (crates/minilake-exec/src/operators/limit.rs:26-117): compacted
// batches in arrival order until offset + limit rows are held.
struct ml_limit {
  ml_ctx* ctx = nullptr;
  int64_t limit = 0, offset = 0;
  std::mutex mu;
  std::atomic<bool> done{false};
  int64_t held = 0;
  std::vector<std::unique_ptr<ml_result>> parts;
};

extern "C" {

ml_status ml_hashagg_create_local(ml_hashagg* global, ml_ctx* ctx, ml_hashagg** out) {
  return guarded(ctx, [&] {
    std::vector<ml_aggregate> aggs;
    ml_agg_spec spec;
    agg_spec_of(global, aggs, spec);
    const ml_status rc = ml_hashagg_create(ctx, &spec, out);
    if (rc != ML_OK) fail(rc, ctx->last_error);
  });
}

ml_status ml_hashagg_combine(ml_hashagg* global, ml_hashagg* local) {
  std::unique_ptr<ml_hashagg> own(local);
  return guarded(global->ctx, [&] {
    if (!local) return;
    std::unique_ptr<ml_result> st(agg_states(local));  // on the local's context (synchronized)
    if (!st) return;
    std::lock_guard<std::mutex> lock(global->combine_mu);
    agg_merge_states(global, st.get());
  });
}

ml_status ml_join_build_create_local(ml_join_build* global, ml_ctx* ctx, ml_join_build** out) {
  return guarded(ctx, [&] {
    ml_join_build_spec spec{};
    spec.keys = global->keys_raw.data();
    spec.key_types = global->key_types.data();
    spec.num_keys = (int32_t)global->key_types.size();
    const ml_status rc = ml_join_build_create(ctx, &spec, out);
    if (rc != ML_OK) fail(rc, ctx->last_error);
  });
}

ml_status ml_join_build_combine(ml_join_build* global, ml_join_build* local) {
  std::unique_ptr<ml_join_build> own(local);
  return guarded(global->ctx, [&] {
    if (!local) return;
    hip_check(hipStreamSynchronize(local->ctx->stream), "synchronize");  // its chunks are complete
    if (!local->schema_fixed) return;
    std::lock_guard<std::mutex> lock(global->combine_mu);
    if (!global->schema_fixed) {
      global->col_types = local->col_types;
      global->col_dicts = local->col_dicts;
      global->key_dicts = local->key_dicts;
      global->col_nullable = local->col_nullable;
      global->schema_fixed = true;
    } else {
      if (global->col_types != local->col_types) fail(ML_ERR_INTERNAL, "join build batches with different schemas");
      for (size_t c = 0; c < local->col_dicts.size(); c++)
        if (!same_dict(global->col_dicts[c], local->col_dicts[c]))
          fail(ML_ERR_UNSUPPORTED, "every build batch must share the string columns' dictionaries");
      for (size_t k = 0; k < local->key_dicts.size(); k++)
        if (!same_dict(global->key_dicts[k], local->key_dicts[k]))
          fail(ML_ERR_UNSUPPORTED, "every build batch must share the key dictionaries");
      for (size_t c = 0; c < local->col_nullable.size() && c < global->col_nullable.size(); c++)
        global->col_nullable[c] = global->col_nullable[c] || local->col_nullable[c];
    }
    for (auto& ch : local->chunks) global->chunks.push_back(std::move(ch));
    local->chunks.clear();
  });
}

ml_status ml_topn_create_local(ml_topn* global, ml_ctx* ctx, ml_topn** out) {
  return guarded(ctx, [&] {
    std::vector<ml_sort_key> keys(global->keys_raw.size());
    for (size_t k = 0; k < keys.size(); k++) {
      keys[k] = ml_sort_key{};
      keys[k].expr = global->keys_raw[k];
      keys[k].asc = global->asc[k];
      keys[k].nulls_first = global->nulls_first[k];
    }
    ml_sort_spec spec{};
    spec.keys = keys.data();
    spec.num_keys = (int32_t)keys.size();
    spec.limit = global->limit;
    const ml_status rc = ml_topn_create(ctx, &spec, out);
    if (rc != ML_OK) fail(rc, ctx->last_error);
  });
}

ml_status ml_topn_combine(ml_topn* global, ml_topn* local) {
  std::unique_ptr<ml_topn> own(local);
  return guarded(global->ctx, [&] {
    if (!local) return;
    hip_check(hipStreamSynchronize(local->ctx->stream), "synchronize");
    if (!local->schema_fixed || local->best.rows == 0) return;
    std::lock_guard<std::mutex> lock(global->combine_mu);
    if (!global->schema_fixed) {
      global->best.types = local->best.types;
      global->best.dicts = local->best.dicts;
      global->best.cols.resize(local->best.types.size());
      global->best.valid.resize(local->best.types.size());
      global->best.rows = 0;
      global->schema_fixed = true;
    } else if (global->best.types != local->best.types) {
      fail(ML_ERR_INTERNAL, "TopN batches with different schemas");
    }
    topn_merge(global, local->best);
    hip_check(hipStreamSynchronize(global->ctx->stream), "synchronize");
  });
}

ml_status ml_limit_create(ml_ctx* ctx, int64_t limit, int64_t offset, ml_limit** out) {
  return guarded(ctx, [&] {
    if (limit < 0 || offset < 0) fail(ML_ERR_PLAN, "negative LIMIT or OFFSET");
    auto* L = new ml_limit();
    L->ctx = ctx;
    L->limit = limit;
    L->offset = offset;
    *out = L;
  });
}

ml_status ml_limit_sink(ml_limit* s, ml_ctx* ctx, const ml_batch* batch) {
  return guarded(ctx, [&] {
    if (s->done.load(std::memory_order_relaxed)) return;
    const int64_t needed = s->limit > INT64_MAX - s->offset ? INT64_MAX : s->limit + s->offset;
    ml_result* r = nullptr;  // Batch::compact on the caller's context
    const ml_status rc = ml_pipeline_collect(ctx, batch, nullptr, 0, &r);
    if (rc != ML_OK) fail(rc, ctx->last_error);
    std::unique_ptr<ml_result> part(r);
    std::lock_guard<std::mutex> lock(s->mu);
    if (s->held >= needed) return;
    s->held += part->num_rows;
    s->parts.push_back(std::move(part));
    if (s->held >= needed) s->done.store(true, std::memory_order_relaxed);
  });
}

int32_t ml_limit_is_finished(const ml_limit* s) { return s && s->done.load(std::memory_order_relaxed) ? 1 : 0; }

ml_status ml_limit_finalize(ml_limit* s, ml_result** out) {
  return guarded(s->ctx, [&] {
    std::lock_guard<std::mutex> lock(s->mu);
    *out = nullptr;
    if (s->parts.empty()) {  // no batch: no rows and no columns
      auto* r = new ml_result();
      r->ctx = s->ctx;
      r->device = s->ctx->device;
      r->finish_views();
      *out = r;
      return;
    }
    std::unique_ptr<ml_result> all(concat_results(s->ctx, s->parts));
    s->parts.clear();
    const int64_t start = std::min(s->offset, all->num_rows);
    const int64_t len = std::min(s->limit, all->num_rows - start);
    *out = slice_result(s->ctx, all.get(), start, len);
  });
}

void ml_limit_free(ml_limit* s) { delete s; }

}  // extern "C"
