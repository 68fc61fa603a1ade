// Batched fragments: many independent
//   SELECT g, SUM(m), COUNT(*) FROM probe JOIN build USING (key) GROUP BY g
// fragments run by two persistent kernels (build, then probe) that pull
// (fragment, tile) work items from one device-side atomic queue, the GPU form
// of MiniLake's morsel queue (crates/minilake-exec/src/scheduler.rs:44-68).
// Small group domains aggregate in LDS per work item and flush once; large
// ones use global FP64 atomics. The whole run can be captured in a hipGraph.
#include <algorithm>
#include <chrono>

#include "engine.h"

namespace mlg {

struct FragDesc {
  const int64_t* bkeys;
  const int64_t* bgroups;
  int64_t brows;
  const int64_t* pkeys;
  const double* pvals;
  int64_t prows;
  int64_t groups;
  double* out_sums;
  int64_t* out_counts;
  int64_t* slots;    // [cap][2]: key, group
  uint64_t mask;
};

constexpr int64_t FRAG_EMPTY = INT64_MIN;  // key value reserved for empty slots

__device__ __forceinline__ int find_item(const uint64_t* start, int nfrags, uint64_t item) {
  int lo = 0, hi = nfrags;  // largest f with start[f] <= item
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (start[mid] <= item) lo = mid;
    else hi = mid;
  }
  return lo;
}

__global__ void frag_reset_kernel(const FragDesc* frags, int nfrags, uint32_t* queues) {
  // zero every output (one workgroup per fragment) and both work queues
  const FragDesc& F = frags[blockIdx.x];
  for (int64_t g = threadIdx.x; g < F.groups; g += blockDim.x) {
    F.out_sums[g] = 0.0;
    F.out_counts[g] = 0;
  }
  for (int64_t s = threadIdx.x; s <= (int64_t)F.mask; s += blockDim.x) F.slots[2 * s] = FRAG_EMPTY;
  if (blockIdx.x == 0 && threadIdx.x < 2) queues[threadIdx.x] = 0;
}

__global__ void __launch_bounds__(1024) frag_build_kernel(const FragDesc* frags, int nfrags, const uint64_t* start,
                                                          uint64_t nitems, uint32_t* queue, int tile_rows) {
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
    const int64_t e = b + tile_rows < F.brows ? b + tile_rows : F.brows;
    for (int64_t r = b + threadIdx.x; r < e; r += blockDim.x) {
      const int64_t k = F.bkeys[r];
      uint64_t pos = hash_u64((uint64_t)k) & F.mask;
      for (;;) {
        unsigned long long* slot = (unsigned long long*)&F.slots[2 * pos];
        if (atomicCAS(slot, (unsigned long long)FRAG_EMPTY, (unsigned long long)k) == (unsigned long long)FRAG_EMPTY) {
          F.slots[2 * pos + 1] = F.bgroups[r];
          break;
        }
        pos = (pos + 1) & F.mask;
      }
    }
  }
}

__global__ void __launch_bounds__(1024) frag_probe_kernel(const FragDesc* frags, int nfrags, const uint64_t* start,
                                                          uint64_t nitems, uint32_t* queue, int tile_rows,
                                                          int lds_groups) {
  extern __shared__ double lds_sum[];  // lds_groups sums, then lds_groups counts
  unsigned long long* lds_cnt = (unsigned long long*)(lds_sum + lds_groups);
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
    const int64_t e = b + tile_rows < F.prows ? b + tile_rows : F.prows;
    const bool local = F.groups <= lds_groups;
    if (local) {
      for (int64_t g = threadIdx.x; g < F.groups; g += blockDim.x) {
        lds_sum[g] = 0.0;
        lds_cnt[g] = 0;
      }
      __syncthreads();
    }
    for (int64_t r = b + threadIdx.x; r < e; r += blockDim.x) {
      const int64_t k = F.pkeys[r];
      uint64_t pos = hash_u64((uint64_t)k) & F.mask;
      int64_t g = -1;
      for (;;) {
        const int64_t sk = F.slots[2 * pos];
        if (sk == k) {
          g = F.slots[2 * pos + 1];
          break;
        }
        if (sk == FRAG_EMPTY) break;
        pos = (pos + 1) & F.mask;
      }
      if (g < 0) continue;
      const double v = F.pvals[r];
      if (local) {
        unsafeAtomicAdd(&lds_sum[g], v);
        atomicAdd(&lds_cnt[g], 1ull);
      } else {
        unsafeAtomicAdd(&F.out_sums[g], v);
        atomicAdd((unsigned long long*)&F.out_counts[g], 1ull);
      }
    }
    if (local) {
      __syncthreads();
      for (int64_t g = threadIdx.x; g < F.groups; g += blockDim.x) {
        if (lds_cnt[g]) {
          unsafeAtomicAdd(&F.out_sums[g], lds_sum[g]);
          atomicAdd((unsigned long long*)&F.out_counts[g], lds_cnt[g]);
        }
      }
    }
    __syncthreads();
  }
}

}  // namespace mlg

struct ml_fragments {
  ml_ctx* ctx = nullptr;
  std::vector<mlg::FragDesc> host;
  mlg::Buffer descs, tables, starts_build, starts_probe, queues;
  uint64_t items_build = 0, items_probe = 0;
  ml_fragment_tuning tune{};
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<mlg::Buffer> per_frag_desc;
};

using namespace mlg;

template <class F>
static ml_status frag_guard(ml_ctx* ctx, F f) {
  try {
    f();
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

static void frag_plan(ml_fragments* F) {
  ml_ctx* ctx = F->ctx;
  const int n = (int)F->host.size();
  std::vector<uint64_t> sb(n + 1, 0), sp(n + 1, 0);
  const int64_t t = F->tune.tile_rows;
  for (int i = 0; i < n; i++) {
    sb[i + 1] = sb[i] + (uint64_t)((F->host[i].brows + t - 1) / t);
    sp[i + 1] = sp[i] + (uint64_t)((F->host[i].prows + t - 1) / t);
  }
  F->items_build = sb[n];
  F->items_probe = sp[n];
  F->starts_build = Buffer(ctx->arena, sb.size() * 8);
  F->starts_probe = Buffer(ctx->arena, sp.size() * 8);
  hip_check(hipMemcpyAsync(F->starts_build.ptr, sb.data(), sb.size() * 8, hipMemcpyHostToDevice, ctx->stream), "upload");
  hip_check(hipMemcpyAsync(F->starts_probe.ptr, sp.data(), sp.size() * 8, hipMemcpyHostToDevice, ctx->stream), "upload");
  hip_check(hipStreamSynchronize(ctx->stream), "sync");
  if (F->exec) {
    (void)hipGraphExecDestroy(F->exec);
    F->exec = nullptr;
  }
  if (F->graph) {
    (void)hipGraphDestroy(F->graph);
    F->graph = nullptr;
  }
}

static void frag_launch_all(ml_fragments* F, hipStream_t s) {
  const int n = (int)F->host.size();
  const int grid = F->ctx->num_cus * std::max(1, F->tune.blocks_per_cu);
  const int block = F->tune.block_size;
  FragDesc* d = F->descs.as<FragDesc>();
  uint32_t* q = F->queues.as<uint32_t>();
  frag_reset_kernel<<<n, 256, 0, s>>>(d, n, q);
  frag_build_kernel<<<grid, block, 0, s>>>(d, n, F->starts_build.as<uint64_t>(), F->items_build, q, F->tune.tile_rows);
  const size_t lds = (size_t)std::max(1, F->tune.lds_groups) * 16;
  frag_probe_kernel<<<grid, block, lds, s>>>(d, n, F->starts_probe.as<uint64_t>(), F->items_probe, q + 1,
                                             F->tune.tile_rows, F->tune.lds_groups);
}

extern "C" {

ml_status ml_fragments_create(ml_ctx* ctx, const ml_fragment* frags, int32_t n, const ml_fragment_tuning* tuning,
                              ml_fragments** out) {
  auto* F = new ml_fragments();
  F->ctx = ctx;
  ml_status st = frag_guard(ctx, [&] {
    if (n <= 0) fail(ML_ERR_PLAN, "no fragments");
    F->tune = tuning ? *tuning : ml_fragment_tuning{256, 4096, 4, 2048};
    if (F->tune.block_size <= 0 || F->tune.block_size > 1024 || F->tune.block_size % 64) F->tune.block_size = 256;
    if (F->tune.tile_rows <= 0) F->tune.tile_rows = 4096;
    if (F->tune.blocks_per_cu <= 0) F->tune.blocks_per_cu = 4;
    if (F->tune.lds_groups < 0 || F->tune.lds_groups > 3968) F->tune.lds_groups = 2048;
    uint64_t total_slots = 0;
    for (int i = 0; i < n; i++) total_slots += pow2_at_least(2 * (uint64_t)std::max<int64_t>(frags[i].build_rows, 1));
    F->tables = Buffer(ctx->arena, total_slots * 16);
    uint64_t off = 0;
    for (int i = 0; i < n; i++) {
      const ml_fragment& m = frags[i];
      if (m.build_rows < 0 || m.probe_rows < 0 || m.num_groups <= 0) fail(ML_ERR_PLAN, "bad fragment sizes");
      FragDesc d{};
      d.bkeys = (const int64_t*)m.build_keys;
      d.bgroups = (const int64_t*)m.build_groups;
      d.brows = m.build_rows;
      d.pkeys = (const int64_t*)m.probe_keys;
      d.pvals = (const double*)m.probe_values;
      d.prows = m.probe_rows;
      d.groups = m.num_groups;
      d.out_sums = (double*)m.out_sums;
      d.out_counts = (int64_t*)m.out_counts;
      const uint64_t cap = pow2_at_least(2 * (uint64_t)std::max<int64_t>(m.build_rows, 1));
      d.slots = F->tables.as<int64_t>() + 2 * off;
      d.mask = cap - 1;
      off += cap;
      F->host.push_back(d);
    }
    F->descs = Buffer(ctx->arena, F->host.size() * sizeof(FragDesc));
    hip_check(hipMemcpyAsync(F->descs.ptr, F->host.data(), F->host.size() * sizeof(FragDesc), hipMemcpyHostToDevice,
                             ctx->stream), "upload");
    F->queues = Buffer(ctx->arena, 256);
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
    hip_check(hipGetLastError(), "fragment launch");
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
  });
}

ml_status ml_fragments_run_per_fragment(ml_fragments* F) {
  return frag_guard(F->ctx, [&] {
    hipStream_t s = F->ctx->stream;
    const int n = (int)F->host.size();
    if (F->per_frag_desc.empty()) {
      // one descriptor + item table + queue pair per fragment
      for (int i = 0; i < n; i++) {
        Buffer b(F->ctx->arena, sizeof(FragDesc) + 64);
        hip_check(hipMemcpyAsync(b.ptr, &F->host[i], sizeof(FragDesc), hipMemcpyHostToDevice, s), "upload");
        F->per_frag_desc.push_back(std::move(b));
      }
      hip_check(hipStreamSynchronize(s), "sync");
    }
    const int64_t t = F->tune.tile_rows;
    std::vector<uint64_t> st(2 * n, 0);
    Buffer starts(F->ctx->arena, 16);
    launch_fill_u64(starts.as<uint64_t>(), 0, 2, s);
    for (int i = 0; i < n; i++) {
      const FragDesc& d = F->host[i];
      FragDesc* dd = F->per_frag_desc[i].as<FragDesc>();
      uint32_t* q = (uint32_t*)(F->per_frag_desc[i].as<uint8_t>() + sizeof(FragDesc) + 32);
      const uint64_t nb = (uint64_t)((d.brows + t - 1) / t), np = (uint64_t)((d.prows + t - 1) / t);
      const int gb = (int)std::max<uint64_t>(1, std::min<uint64_t>(nb, (uint64_t)F->ctx->num_cus * 4));
      const int gp = (int)std::max<uint64_t>(1, std::min<uint64_t>(np, (uint64_t)F->ctx->num_cus * 4));
      frag_reset_kernel<<<1, 256, 0, s>>>(dd, 1, q);
      frag_build_kernel<<<gb, F->tune.block_size, 0, s>>>(dd, 1, starts.as<uint64_t>(), nb, q, (int)t);
      frag_probe_kernel<<<gp, F->tune.block_size, (size_t)std::max(1, F->tune.lds_groups) * 16, s>>>(
          dd, 1, starts.as<uint64_t>(), np, q + 1, (int)t, F->tune.lds_groups);
    }
    hip_check(hipGetLastError(), "fragment launch");
  });
}

ml_status ml_fragments_autotune(ml_fragments* F, ml_fragment_tuning* best) {
  return frag_guard(F->ctx, [&] {
    hipStream_t s = F->ctx->stream;
    const int blocks[] = {256, 512};
    const int tiles[] = {2048, 8192, 32768};
    const int per_cu[] = {2, 4, 8};
    const int lds[] = {0, 1024, 2048};
    hipEvent_t a, b;
    hip_check(hipEventCreate(&a), "event");
    hip_check(hipEventCreate(&b), "event");
    float best_ms = 1e30f;
    ml_fragment_tuning best_t = F->tune;
    for (int bs : blocks)
      for (int tr : tiles)
        for (int pc : per_cu)
          for (int lg : lds) {
            F->tune = ml_fragment_tuning{bs, tr, pc, lg};
            frag_plan(F);
            frag_launch_all(F, s);  // warm
            hip_check(hipEventRecord(a, s), "event");
            for (int r = 0; r < 2; r++) frag_launch_all(F, s);
            hip_check(hipEventRecord(b, s), "event");
            hip_check(hipEventSynchronize(b), "event");
            float ms = 0;
            hip_check(hipEventElapsedTime(&ms, a, b), "event");
            if (ms < best_ms) {
              best_ms = ms;
              best_t = F->tune;
            }
          }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    F->tune = best_t;
    frag_plan(F);
    if (best) *best = best_t;
  });
}

void ml_fragments_free(ml_fragments* F) {
  if (!F) return;
  if (F->exec) (void)hipGraphExecDestroy(F->exec);
  if (F->graph) (void)hipGraphDestroy(F->graph);
  delete F;
}

}  // extern "C"
