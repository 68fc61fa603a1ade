// Batched fragments: many independent
//   SELECT g, SUM(v), COUNT(*) FROM probe JOIN build USING (key) GROUP BY g
// fragments (g = the group id carried by each build row) run by persistent
// kernels that pull (fragment, tile) work items from device-side atomic
// queues, the GPU form of MiniLake's morsel queue
// (crates/minilake-exec/src/scheduler.rs:44-68). A probe row contributes to
// every build row with an equal key, as HashJoinProbe emits every chain match
// (crates/minilake-exec/src/operators/join.rs:283-300); NULL keys never match.
//
// Per fragment, an open-addressed table of 16-byte slots {key, (group << 1) | 1}
// (second word 0 = empty; no key value is reserved) at load <= 1/2, linear
// probing that always runs to the first empty slot, so duplicate keys need no
// separate chains. Groups accumulate in LDS: the whole domain as an array when
// it fits, otherwise an LDS cache (4 probes from g mod entries) in front of
// FP64 / u64 global atomics, so a hot group of a skewed probe side is added
// up on chip and flushed once per work item.
//
// Launch sequence (capturable in a hipGraph; no memset nodes): reset (tables,
// outputs, queues, the error words), build, probe, and a copy of the error
// words to pinned host memory that the run functions check.
#include <algorithm>
#include <chrono>

#include "engine.h"

namespace mlg {

struct FragDesc {
  const int64_t* bkeys;
  const int64_t* bgroups;
  const uint64_t* bvalid;  // LSB-first bitmap or null
  int64_t brows;
  const int64_t* pkeys;
  const double* pvals;
  const uint64_t* pvalid;
  int64_t prows;
  int64_t groups;
  double* out_sums;
  int64_t* out_counts;
  uint64_t* slots;         // [cap][2]
  uint64_t mask;           // cap - 1
  int32_t agg;             // FRAG_AGG_*
  int32_t _pad;
  uint64_t zero_start;     // first zeroing word of this fragment (reset kernel)
};

enum : int32_t { FRAG_AGG_GLOBAL = 0, FRAG_AGG_ARRAY = 1, FRAG_AGG_CACHE = 2 };

// err[0]: a build group id outside [0, groups) was seen (its row skipped);
// err[1]: the lowest fragment index with such a row.
constexpr int ERR_WORDS = 2;
constexpr int RESET_ITEM_WORDS = 4096;  // u64 words zeroed per reset work item
constexpr int CACHE_PROBES = 4;

__device__ __forceinline__ int find_item(const uint64_t* start, int nfrags, uint64_t item) {
  int lo = 0, hi = nfrags;  // largest f with start[f] <= item
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (start[mid] <= item) lo = mid;
    else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ uint64_t frag_hash(uint64_t k) { return tmix(k ^ TSEED); }

__device__ __forceinline__ bool bit_valid(const uint64_t* bits, int64_t r) {
  return !bits || ((bits[r >> 6] >> (r & 63)) & 1ull);
}

// Zero every fragment's table slots, sums and counts: item i covers words
// [i * RESET_ITEM_WORDS, ...) of the concatenation over fragments of
// (2 * cap table words, groups sum words, groups count words).
__global__ void __launch_bounds__(256) frag_reset_kernel(const FragDesc* frags, int nfrags, uint64_t total_words,
                                                         uint32_t* queues, uint32_t* err) {
  if (blockIdx.x == 0 && threadIdx.x < 2) queues[threadIdx.x] = 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    err[0] = 0;
    err[1] = 0x7FFFFFFFu;
  }
  const uint64_t nitems = (total_words + RESET_ITEM_WORDS - 1) / RESET_ITEM_WORDS;
  for (uint64_t item = blockIdx.x; item < nitems; item += gridDim.x) {
    const uint64_t w0 = item * RESET_ITEM_WORDS;
    const uint64_t w1 = min(total_words, w0 + RESET_ITEM_WORDS);
    // fragments whose range intersects [w0, w1): found by binary search, then walked
    int lo = 0, hi = nfrags;
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (frags[mid].zero_start <= w0) lo = mid;
      else hi = mid;
    }
    for (int f = lo; f < nfrags && frags[f].zero_start < w1; f++) {
      const FragDesc& F = frags[f];
      const uint64_t tw = 2 * (F.mask + 1), gw = (uint64_t)F.groups;
      const uint64_t fb = F.zero_start, fe = fb + tw + 2 * gw;
      const uint64_t a = max(w0, fb), b = min(w1, fe);
      for (uint64_t w = a + threadIdx.x; w < b; w += blockDim.x) {
        const uint64_t o = w - fb;
        if (o < tw) F.slots[o] = 0;
        else if (o < tw + gw) F.out_sums[o - tw] = 0.0;
        else F.out_counts[o - tw - gw] = 0;
      }
    }
  }
}

__global__ void __launch_bounds__(1024) frag_build_kernel(const FragDesc* frags, int nfrags, const uint64_t* start,
                                                          uint64_t nitems, uint32_t* queue, int tile_rows,
                                                          uint32_t* err, int frag_base) {
  __shared__ uint64_t item_s;
  for (;;) {
    if (threadIdx.x == 0) item_s = atomicAdd(queue, 1u);
    __syncthreads();
    const uint64_t item = item_s;
    __syncthreads();
    if (item >= nitems) break;
    const int f = find_item(start, nfrags, item);
    const FragDesc& F = frags[f];
    const int64_t b = (int64_t)(item - start[f]) * tile_rows;
    const int64_t e = min(b + (int64_t)tile_rows, F.brows);
    for (int64_t r = b + threadIdx.x; r < e; r += blockDim.x) {
      if (!bit_valid(F.bvalid, r)) continue;  // a NULL key is never found (join.rs:180-184)
      const int64_t g = F.bgroups[r];
      if (g < 0 || g >= F.groups) {  // skipped, so no write lands outside the outputs
        atomicOr(&err[0], 1u);
        atomicMin(&err[1], (uint32_t)(frag_base + f));
        continue;
      }
      const uint64_t k = (uint64_t)F.bkeys[r];
      uint64_t pos = frag_hash(k) & F.mask;
      // load <= 1/2: an empty slot is always reached
      for (;;) {
        unsigned long long* w = (unsigned long long*)&F.slots[2 * pos + 1];
        if (atomicCAS(w, 0ull, ((unsigned long long)g << 1) | 1ull) == 0ull) {
          F.slots[2 * pos] = k;
          break;
        }
        pos = (pos + 1) & F.mask;
      }
    }
  }
}

// One matched pair: add v to group g.
__device__ __forceinline__ void frag_add(const FragDesc& F, int agg, int lds_n, double* ls, unsigned long long* lc,
                                         uint32_t* tags, uint64_t g, double v) {
  if (agg == FRAG_AGG_ARRAY) {
    unsafeAtomicAdd(&ls[g], v);
    atomicAdd(&lc[g], 1ull);
    return;
  }
  if (agg == FRAG_AGG_CACHE) {
    const uint32_t want = (uint32_t)g + 1u;
    uint32_t ci = (uint32_t)g & (uint32_t)(lds_n - 1);
#pragma unroll
    for (int pr = 0; pr < CACHE_PROBES; pr++) {
      uint32_t tag = tags[ci];
      if (tag == 0) {
        const uint32_t prev = atomicCAS(&tags[ci], 0u, want);
        tag = prev == 0 ? want : prev;
      }
      if (tag == want) {
        unsafeAtomicAdd(&ls[ci], v);
        atomicAdd(&lc[ci], 1ull);
        return;
      }
      ci = (ci + 1u) & (uint32_t)(lds_n - 1);
    }
  }
  unsafeAtomicAdd(&F.out_sums[g], v);
  atomicAdd((unsigned long long*)&F.out_counts[g], 1ull);
}

// Probe kernel, PW probe rows per lane in flight: the key, value and first
// slot of all PW rows are requested before any is used, then the rows walk
// their probe sequences in lockstep (one 16-byte slot load per row and step)
// until each reaches an empty slot.
template <int PW>
__global__ void __launch_bounds__(1024) frag_probe_kernel(const FragDesc* frags, int nfrags, const uint64_t* start,
                                                          uint64_t nitems, uint32_t* queue, int tile_rows, int lds_n) {
  extern __shared__ uint64_t lds_raw[];
  double* ls = (double*)lds_raw;                                         // lds_n sums
  unsigned long long* lc = (unsigned long long*)(lds_raw + lds_n);       // lds_n counts
  uint32_t* tags = (uint32_t*)(lds_raw + 2 * (size_t)lds_n);             // lds_n tags (cache mode)
  __shared__ uint64_t item_s;
  for (;;) {
    if (threadIdx.x == 0) item_s = atomicAdd(queue, 1u);
    __syncthreads();
    const uint64_t item = item_s;
    __syncthreads();
    if (item >= nitems) break;
    const int f = find_item(start, nfrags, item);
    const FragDesc& F = frags[f];
    const int agg = F.agg;
    const int64_t b = (int64_t)(item - start[f]) * tile_rows;
    const int64_t e = min(b + (int64_t)tile_rows, F.prows);
    const int nloc = agg == FRAG_AGG_ARRAY ? (int)F.groups : agg == FRAG_AGG_CACHE ? lds_n : 0;
    for (int i = threadIdx.x; i < nloc; i += blockDim.x) {
      ls[i] = 0.0;
      lc[i] = 0;
      if (agg == FRAG_AGG_CACHE) tags[i] = 0;
    }
    __syncthreads();
    const uint64_t* slots = F.slots;
    const uint64_t mask = F.mask;
    for (int64_t r0 = b; r0 < e; r0 += (int64_t)PW * blockDim.x) {
      uint64_t k[PW], pos[PW];
      double v[PW];
      bool pend[PW];
#pragma unroll
      for (int u = 0; u < PW; u++) {
        const int64_t r = r0 + (int64_t)u * blockDim.x + threadIdx.x;
        pend[u] = r < e;
        const int64_t rr = pend[u] ? r : b;
        k[u] = (uint64_t)__builtin_nontemporal_load(F.pkeys + rr);
        v[u] = __builtin_nontemporal_load(F.pvals + rr);
        pend[u] = pend[u] && bit_valid(F.pvalid, rr);
        pos[u] = frag_hash(k[u]) & mask;
      }
      for (;;) {
        bool any = false;
#pragma unroll
        for (int u = 0; u < PW; u++) any = any || pend[u];
        if (!any) break;
        ulonglong2 s[PW];
#pragma unroll
        for (int u = 0; u < PW; u++) s[u] = *(const ulonglong2*)(slots + 2 * (pend[u] ? pos[u] : 0));
#pragma unroll
        for (int u = 0; u < PW; u++) {
          if (!pend[u]) continue;
          if (s[u].y == 0) {
            pend[u] = false;
            continue;
          }
          if (s[u].x == k[u]) frag_add(F, agg, lds_n, ls, lc, tags, s[u].y >> 1, v[u]);
          pos[u] = (pos[u] + 1) & mask;
        }
      }
    }
    if (nloc) {
      __syncthreads();
      for (int i = threadIdx.x; i < nloc; i += blockDim.x) {
        if (!lc[i]) continue;
        const uint64_t g = agg == FRAG_AGG_ARRAY ? (uint64_t)i : (uint64_t)(tags[i] - 1u);
        unsafeAtomicAdd(&F.out_sums[g], ls[i]);
        atomicAdd((unsigned long long*)&F.out_counts[g], lc[i]);
      }
    }
    __syncthreads();
  }
}

}  // namespace mlg

struct ml_fragments {
  ml_ctx* ctx = nullptr;
  std::vector<mlg::FragDesc> host;
  mlg::Buffer descs, tables, starts_build, starts_probe, queues, err;
  uint64_t items_build = 0, items_probe = 0, zero_words = 0;
  ml_fragment_tuning tune{};
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  uint32_t* err_host = nullptr;  // pinned: the error words of the last run
  std::vector<mlg::Buffer> per_frag_desc, per_frag_starts;
  ~ml_fragments() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (err_host) (void)hipHostFree(err_host);
  }
};

using namespace mlg;

template <class F>
static ml_status frag_guard(ml_ctx* ctx, F f) {
  try {
    f();
    if (ctx) ctx->last_error.clear();
    return ML_OK;
  } catch (const Error& e) {
    if (ctx) ctx->last_error = e.what();
    return e.code;
  } catch (const std::exception& e) {
    if (ctx) ctx->last_error = e.what();
    return ML_ERR_INTERNAL;
  }
}

static uint64_t pow2_at_least(uint64_t x) {
  uint64_t p = 16;
  while (p < x) p <<= 1;
  return p;
}

static const ml_fragment_tuning kDefaultTuning{256, 8192, 4, 2048, 4, 1};

static void frag_normalize(ml_fragment_tuning& t) {
  if (t.block_size != 256 && t.block_size != 512 && t.block_size != 1024) t.block_size = 256;
  if (t.tile_rows <= 0) t.tile_rows = 8192;
  if (t.blocks_per_cu <= 0 || t.blocks_per_cu > 16) t.blocks_per_cu = 4;
  // LDS entries: a power of two, at most 2048 (2048 x 20 B = 40 KB)
  if (t.lds_groups < 0 || t.lds_groups > 2048) t.lds_groups = 2048;
  if (t.lds_groups > 0) {
    int p = 1;
    while (p * 2 <= t.lds_groups) p *= 2;
    t.lds_groups = p;
  }
  if (t.probe_width != 1 && t.probe_width != 2 && t.probe_width != 4 && t.probe_width != 8) t.probe_width = 4;
  t.lds_cache = t.lds_cache ? 1 : 0;
}

// Work item tables, aggregation modes and descriptors for the current tuning.
static void frag_plan(ml_fragments* F) {
  ml_ctx* ctx = F->ctx;
  const int n = (int)F->host.size();
  std::vector<uint64_t> sb(n + 1, 0), sp(n + 1, 0);
  const int64_t t = F->tune.tile_rows;
  for (int i = 0; i < n; i++) {
    FragDesc& d = F->host[i];
    sb[i + 1] = sb[i] + (uint64_t)((d.brows + t - 1) / t);
    sp[i + 1] = sp[i] + (uint64_t)((d.prows + t - 1) / t);
    if (F->tune.lds_groups > 0 && d.groups <= F->tune.lds_groups) d.agg = FRAG_AGG_ARRAY;
    else if (F->tune.lds_groups > 0 && F->tune.lds_cache && d.groups < (int64_t)UINT32_MAX) d.agg = FRAG_AGG_CACHE;
    else d.agg = FRAG_AGG_GLOBAL;
  }
  F->items_build = sb[n];
  F->items_probe = sp[n];
  if (F->items_build >= UINT32_MAX || F->items_probe >= UINT32_MAX) fail(ML_ERR_UNSUPPORTED, "too many work items");
  F->starts_build = Buffer(ctx->arena, sb.size() * 8);
  F->starts_probe = Buffer(ctx->arena, sp.size() * 8);
  hip_check(hipMemcpyAsync(F->starts_build.ptr, sb.data(), sb.size() * 8, hipMemcpyHostToDevice, ctx->stream), "upload");
  hip_check(hipMemcpyAsync(F->starts_probe.ptr, sp.data(), sp.size() * 8, hipMemcpyHostToDevice, ctx->stream), "upload");
  hip_check(hipMemcpyAsync(F->descs.ptr, F->host.data(), F->host.size() * sizeof(FragDesc), hipMemcpyHostToDevice,
                           ctx->stream), "upload");
  hip_check(hipStreamSynchronize(ctx->stream), "sync");
  if (F->exec) {
    (void)hipGraphExecDestroy(F->exec);
    F->exec = nullptr;
  }
  if (F->graph) {
    (void)hipGraphDestroy(F->graph);
    F->graph = nullptr;
  }
  F->per_frag_desc.clear();
  F->per_frag_starts.clear();
}

static size_t probe_lds(const ml_fragment_tuning& t) { return (size_t)std::max(1, t.lds_groups) * 20; }

static void launch_probe(const ml_fragment_tuning& t, int grid, const FragDesc* d, int n, const uint64_t* starts,
                         uint64_t nitems, uint32_t* q, hipStream_t s) {
  const size_t lds = probe_lds(t);
  const int ln = std::max(1, t.lds_groups);
  switch (t.probe_width) {
    case 1: frag_probe_kernel<1><<<grid, t.block_size, lds, s>>>(d, n, starts, nitems, q, t.tile_rows, ln); break;
    case 2: frag_probe_kernel<2><<<grid, t.block_size, lds, s>>>(d, n, starts, nitems, q, t.tile_rows, ln); break;
    case 8: frag_probe_kernel<8><<<grid, t.block_size, lds, s>>>(d, n, starts, nitems, q, t.tile_rows, ln); break;
    default: frag_probe_kernel<4><<<grid, t.block_size, lds, s>>>(d, n, starts, nitems, q, t.tile_rows, ln); break;
  }
}

static int reset_grid(uint64_t words, int num_cus) {
  const uint64_t items = (words + RESET_ITEM_WORDS - 1) / RESET_ITEM_WORDS;
  return (int)std::max<uint64_t>(1, std::min<uint64_t>(items, (uint64_t)num_cus * 8));
}

static void frag_launch_all(ml_fragments* F, hipStream_t s) {
  const int n = (int)F->host.size();
  const int grid = F->ctx->num_cus * std::max(1, F->tune.blocks_per_cu);
  FragDesc* d = F->descs.as<FragDesc>();
  uint32_t* q = F->queues.as<uint32_t>();
  uint32_t* err = F->err.as<uint32_t>();
  frag_reset_kernel<<<reset_grid(F->zero_words, F->ctx->num_cus), 256, 0, s>>>(d, n, F->zero_words, q, err);
  frag_build_kernel<<<grid, F->tune.block_size, 0, s>>>(d, n, F->starts_build.as<uint64_t>(), F->items_build, q,
                                                        F->tune.tile_rows, err, 0);
  launch_probe(F->tune, grid, d, n, F->starts_probe.as<uint64_t>(), F->items_probe, q + 1, s);
  hip_check(hipMemcpyAsync(F->err_host, err, ERR_WORDS * 4, hipMemcpyDeviceToHost, s), "copy to host");
}

// Wait for the run and turn its error words into a status.
static void frag_finish(ml_fragments* F) {
  hip_check(hipGetLastError(), "fragment launch");
  hip_check(hipStreamSynchronize(F->ctx->stream), "synchronize");
  if (F->err_host[0])
    fail(ML_ERR_PLAN, "fragment " + std::to_string(F->err_host[1]) +
                          ": a build group id lies outside [0, num_groups); such rows were skipped");
}

extern "C" {

ml_status ml_fragments_create(ml_ctx* ctx, const ml_fragment* frags, int32_t n, const ml_fragment_tuning* tuning,
                              ml_fragments** out) {
  auto* F = new ml_fragments();
  F->ctx = ctx;
  ml_status st = frag_guard(ctx, [&] {
    if (!frags || n <= 0) fail(ML_ERR_PLAN, "no fragments");
    F->tune = tuning ? *tuning : kDefaultTuning;
    frag_normalize(F->tune);
    uint64_t total_slots = 0;
    for (int i = 0; i < n; i++) {
      const ml_fragment& m = frags[i];
      if (m.build_rows < 0 || m.probe_rows < 0 || m.num_groups <= 0)
        fail(ML_ERR_PLAN, "fragment " + std::to_string(i) + ": bad sizes");
      if ((m.build_rows > 0 && (!m.build_keys || !m.build_groups)) ||
          (m.probe_rows > 0 && (!m.probe_keys || !m.probe_values)) || !m.out_sums || !m.out_counts)
        fail(ML_ERR_PLAN, "fragment " + std::to_string(i) + ": null column");
      if (m.num_groups > ((int64_t)1 << 62)) fail(ML_ERR_UNSUPPORTED, "fragment group domain beyond 2^62");
      total_slots += pow2_at_least(2 * (uint64_t)std::max<int64_t>(m.build_rows, 1));
    }
    F->tables = Buffer(ctx->arena, total_slots * 16);
    uint64_t off = 0, zw = 0;
    for (int i = 0; i < n; i++) {
      const ml_fragment& m = frags[i];
      FragDesc d{};
      d.bkeys = m.build_keys;
      d.bgroups = m.build_groups;
      d.bvalid = m.build_validity;
      d.brows = m.build_rows;
      d.pkeys = m.probe_keys;
      d.pvals = m.probe_values;
      d.pvalid = m.probe_validity;
      d.prows = m.probe_rows;
      d.groups = m.num_groups;
      d.out_sums = m.out_sums;
      d.out_counts = m.out_counts;
      const uint64_t cap = pow2_at_least(2 * (uint64_t)std::max<int64_t>(m.build_rows, 1));
      d.slots = F->tables.as<uint64_t>() + 2 * off;
      d.mask = cap - 1;
      d.zero_start = zw;
      off += cap;
      zw += 2 * cap + 2 * (uint64_t)m.num_groups;
      F->host.push_back(d);
    }
    F->zero_words = zw;
    F->descs = Buffer(ctx->arena, F->host.size() * sizeof(FragDesc));
    F->queues = Buffer(ctx->arena, 256);
    F->err = Buffer(ctx->arena, 256);
    hip_check(hipHostMalloc((void**)&F->err_host, 256, hipHostMallocDefault), "hipHostMalloc");
    frag_plan(F);
  });
  if (st != ML_OK) {
    delete F;
    *out = nullptr;
    return st;
  }
  *out = F;
  return ML_OK;
}

ml_status ml_fragments_run(ml_fragments* F) {
  return frag_guard(F->ctx, [&] {
    frag_launch_all(F, F->ctx->stream);
    frag_finish(F);
  });
}

ml_status ml_fragments_run_graph(ml_fragments* F) {
  return frag_guard(F->ctx, [&] {
    hipStream_t s = F->ctx->stream;
    if (!F->exec) {
      hip_check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
      frag_launch_all(F, s);
      hip_check(hipStreamEndCapture(s, &F->graph), "hipStreamEndCapture");
      hip_check(hipGraphInstantiate(&F->exec, F->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
    }
    hip_check(hipGraphLaunch(F->exec, s), "hipGraphLaunch");
    frag_finish(F);
  });
}

ml_status ml_fragments_run_per_fragment(ml_fragments* F) {
  return frag_guard(F->ctx, [&] {
    hipStream_t s = F->ctx->stream;
    const int n = (int)F->host.size();
    const int64_t t = F->tune.tile_rows;
    if (F->per_frag_desc.empty()) {
      // per fragment: its descriptor (zeroing offset 0) and item tables {0, items}
      for (int i = 0; i < n; i++) {
        FragDesc d = F->host[i];
        d.zero_start = 0;
        Buffer b(F->ctx->arena, sizeof(FragDesc));
        hip_check(hipMemcpyAsync(b.ptr, &d, sizeof(FragDesc), hipMemcpyHostToDevice, s), "upload");
        const uint64_t st[4] = {0, (uint64_t)((d.brows + t - 1) / t), 0, (uint64_t)((d.prows + t - 1) / t)};
        Buffer sb(F->ctx->arena, sizeof(st));
        hip_check(hipMemcpyAsync(sb.ptr, st, sizeof(st), hipMemcpyHostToDevice, s), "upload");
        hip_check(hipStreamSynchronize(s), "sync");
        F->per_frag_desc.push_back(std::move(b));
        F->per_frag_starts.push_back(std::move(sb));
      }
    }
    uint32_t* q = F->queues.as<uint32_t>();
    uint32_t* err = F->err.as<uint32_t>();
    for (int i = 0; i < n; i++) {
      const FragDesc& d = F->host[i];
      const FragDesc* dd = F->per_frag_desc[i].as<FragDesc>();
      const uint64_t* st = F->per_frag_starts[i].as<uint64_t>();
      const uint64_t nb = (uint64_t)((d.brows + t - 1) / t), np = (uint64_t)((d.prows + t - 1) / t);
      const int cap_grid = F->ctx->num_cus * std::max(1, F->tune.blocks_per_cu);
      const int gb = (int)std::max<uint64_t>(1, std::min<uint64_t>(nb, (uint64_t)cap_grid));
      const int gp = (int)std::max<uint64_t>(1, std::min<uint64_t>(np, (uint64_t)cap_grid));
      const uint64_t words = 2 * (d.mask + 1) + 2 * (uint64_t)d.groups;
      // the error words are cleared by the first fragment's reset only
      uint32_t* e = err + (i == 0 ? 0 : 2);
      frag_reset_kernel<<<reset_grid(words, F->ctx->num_cus), 256, 0, s>>>(dd, 1, words, q, e);
      frag_build_kernel<<<gb, F->tune.block_size, 0, s>>>(dd, 1, st, nb, q, (int)t, err, i);
      launch_probe(F->tune, gp, dd, 1, st + 2, np, q + 1, s);
    }
    hip_check(hipMemcpyAsync(F->err_host, err, ERR_WORDS * 4, hipMemcpyDeviceToHost, s), "copy to host");
    frag_finish(F);
  });
}

// Coordinate descent over the tuning: each parameter in turn takes the value
// whose timed run (two launches after a warm-up) is fastest with the others
// fixed, until a sweep changes nothing (at most three sweeps).
ml_status ml_fragments_autotune(ml_fragments* F, ml_fragment_tuning* best) {
  return frag_guard(F->ctx, [&] {
    hipStream_t s = F->ctx->stream;
    hipEvent_t a, b;
    hip_check(hipEventCreate(&a), "event");
    hip_check(hipEventCreate(&b), "event");
    auto time_of = [&](const ml_fragment_tuning& t) {
      F->tune = t;
      frag_plan(F);
      frag_launch_all(F, s);  // warm
      frag_finish(F);
      hip_check(hipEventRecord(a, s), "event");
      for (int r = 0; r < 2; r++) frag_launch_all(F, s);
      hip_check(hipEventRecord(b, s), "event");
      frag_finish(F);
      float ms = 0;
      hip_check(hipEventElapsedTime(&ms, a, b), "event");
      return ms;
    };
    ml_fragment_tuning cur = F->tune;
    float cur_ms = time_of(cur);
    const std::vector<int> blocks{256, 512, 1024}, tiles{2048, 8192, 32768, 131072}, per_cu{1, 2, 4, 8},
        lds{0, 512, 2048}, width{1, 2, 4, 8}, cache{0, 1};
    struct Dim {
      int32_t ml_fragment_tuning::*field;
      const std::vector<int>* values;
    };
    const Dim dims[] = {{&ml_fragment_tuning::probe_width, &width}, {&ml_fragment_tuning::tile_rows, &tiles},
                        {&ml_fragment_tuning::block_size, &blocks}, {&ml_fragment_tuning::blocks_per_cu, &per_cu},
                        {&ml_fragment_tuning::lds_groups, &lds},    {&ml_fragment_tuning::lds_cache, &cache}};
    for (int sweep = 0; sweep < 3; sweep++) {
      bool changed = false;
      for (const Dim& dm : dims) {
        for (int v : *dm.values) {
          if (cur.*dm.field == v) continue;
          ml_fragment_tuning t = cur;
          t.*dm.field = v;
          frag_normalize(t);
          if (probe_lds(t) > 64 * 1024) continue;
          const float ms = time_of(t);
          if (ms < cur_ms * 0.98f) {
            cur = t;
            cur_ms = ms;
            changed = true;
          }
        }
      }
      if (!changed) break;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    F->tune = cur;
    frag_plan(F);
    if (best) *best = cur;
  });
}

void ml_fragments_free(ml_fragments* F) { delete F; }

}  // extern "C"
