// Allocation-free filter: ml_filter for callers that cannot allocate device
// memory (code under HIP graph capture, serving loops, hosts that forbid
// hipMalloc). Every device buffer of the call comes from one file-scope
// device array, `scratch`, that each GPU holds until the process ends.
//
// The array's name is only a host-side stand-in for the device symbol, so the
// host never passes it to a kernel: ml_ctx_create calls noalloc_setup once per
// device, which asks HIP for that device's address of the array
// (hipGetSymbolAddress) and keeps it; the entry point uses that address.
#include <mutex>

#include "engine.h"

namespace mlg {

static __device__ __align__(256) uint8_t scratch[ML_FILTER_NOALLOC_SCRATCH_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex noalloc_mu;
static uint8_t* noalloc_ptr[NOALLOC_MAX_DEVICES] = {};

// Called with `device` current, from ml_ctx_create.
void noalloc_setup(int device) {
  if (device < 0 || device >= NOALLOC_MAX_DEVICES) fail(ML_ERR_EXECUTION, "device ordinal beyond the scratch table");
  std::lock_guard<std::mutex> lock(noalloc_mu);
  if (noalloc_ptr[device]) return;
  void* p = nullptr;
  hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
  if (!p) fail(ML_ERR_INTERNAL, "hipGetSymbolAddress(scratch) returned no address");
  noalloc_ptr[device] = (uint8_t*)p;
}

static uint8_t* noalloc_base(int device) {
  std::lock_guard<std::mutex> lock(noalloc_mu);
  uint8_t* p = (device >= 0 && device < NOALLOC_MAX_DEVICES) ? noalloc_ptr[device] : nullptr;
  if (!p) fail(ML_ERR_INTERNAL, "the scratch array of this device was not set up");
  return p;
}

static size_t round256(size_t n) { return (n + 255) & ~(size_t)255; }

}  // namespace mlg

using namespace mlg;

extern "C" ml_status ml_filter_noalloc(ml_ctx* ctx, const ml_expr* predicate, const ml_batch* in, uint32_t* out_sel,
                                       int64_t* out_len) {
  if (!ctx) return ML_ERR_INTERNAL;
  try {
    if (!in || !predicate || !out_len) fail(ML_ERR_INTERNAL, "null argument");
    const int64_t active = in->selection ? in->selection_len : in->num_rows;
    if (active < 0 || in->num_rows < 0) fail(ML_ERR_INTERNAL, "negative row count");
    if (active > ML_FILTER_NOALLOC_MAX_ROWS)
      fail(ML_ERR_RESOURCES_EXHAUSTED, "ml_filter_noalloc takes at most " + std::to_string(ML_FILTER_NOALLOC_MAX_ROWS) +
                                           " active rows, got " + std::to_string(active));
    // what the call takes from the array: one flag byte per active row, the
    // compaction's block counts and offsets, a byte per row for each column
    // with a validity bitmap, and the program and lookup tables (bounded below)
    const int64_t nb = compact_blocks(active);
    size_t need = round256((size_t)std::max<int64_t>(active, 1)) + round256((size_t)std::max<int64_t>(nb, 1) * 4) +
                  round256((size_t)(nb + 1) * 8);
    for (int i = 0; i < in->num_columns; i++)
      if (in->columns[i].validity && in->num_rows > 0) need += round256((size_t)in->num_rows);
    if (need > ML_FILTER_NOALLOC_SCRATCH_BYTES)
      fail(ML_ERR_RESOURCES_EXHAUSTED, "ml_filter_noalloc needs " + std::to_string(need) + " scratch bytes; the array holds " +
                                           std::to_string((size_t)ML_FILTER_NOALLOC_SCRATCH_BYTES));
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    ScratchBump bump{noalloc_base(ctx->device), (size_t)ML_FILTER_NOALLOC_SCRATCH_BYTES, 0};
    // the compaction buffers first, so the plan's own buffers use what is left
    uint8_t* flags = (uint8_t*)bump.take((size_t)std::max<int64_t>(active, 1));
    uint32_t* counts = (uint32_t*)bump.take((size_t)std::max<int64_t>(nb, 1) * 4);
    uint64_t* offsets = (uint64_t*)bump.take((size_t)(nb + 1) * 8);
    PipelineBuild B(ctx);
    B.use_scratch(&bump);
    B.source(in);
    ml_operator op{};
    op.kind = ML_PIPE_FILTER;
    op.predicate = *predicate;
    B.add_ops(&op, 1);
    B.P.sink = SINK_FLAGS;
    B.P.flag_out = flags;
    B.upload();
    if (active > 0) {
      const int block = ctx->cfg.block_size > 0 ? ctx->cfg.block_size : 256;
      launch_pipeline(B.P, B.grid(block), block, 0, ctx->stream);
    }
    // every flag byte of [0, active) is written by the filter, so nothing in
    // the array needs clearing between calls
    launch_compact(flags, active, counts, offsets, B.P.sel, out_sel, 1, ctx->stream);
    hip_check(hipGetLastError(), "filter launch");
    uint64_t total = 0;
    hip_check(hipMemcpyAsync(ctx->pinned, offsets + nb, 8, hipMemcpyDeviceToHost, ctx->stream), "copy to host");
    hip_check(hipStreamSynchronize(ctx->stream), "synchronize");
    std::memcpy(&total, ctx->pinned, 8);
    *out_len = (int64_t)total;
    ctx->last_error.clear();
    return ML_OK;
  } catch (const Error& e) {
    ctx->last_error = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    ctx->last_error = "host allocation failed";
    return ML_ERR_RESOURCES_EXHAUSTED;
  } catch (const std::exception& e) {
    ctx->last_error = e.what();
    return ML_ERR_INTERNAL;
  }
}
