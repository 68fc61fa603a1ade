// Learned cardinality estimation: a 32-64-64-2 FP32 ReLU MLP trained and run
// with hand-written kernels, plus the feature extractor (row-group min/max
// statistics and a sampled KMV sketch of an int64 key column).
//
// Kernel design (wave64): one wave evaluates one sample at a time with lane j
// owning hidden unit j of both layers. Weights sit in LDS in the orientation
// each access needs (consecutive lanes read consecutive words, no bank
// conflicts); activations move between lanes with v_readlane. In the backward
// pass lane j accumulates row j of dW1 and dW2 and column j of dW3 in
// registers over the wave's samples; per-wave partial gradients are then
// summed in a fixed order, so gradients are deterministic run to run.
#include <algorithm>
#include <cmath>

#include "engine.h"

namespace mlg {

constexpr int IN = ML_MLP_IN, H1 = ML_MLP_H1, H2 = ML_MLP_H2, OUT = ML_MLP_OUT;
constexpr int OFF_W1 = 0;                    // W1[H1][IN]
constexpr int OFF_B1 = OFF_W1 + H1 * IN;     // b1[H1]
constexpr int OFF_W2 = OFF_B1 + H1;          // W2[H2][H1]
constexpr int OFF_B2 = OFF_W2 + H2 * H1;     // b2[H2]
constexpr int OFF_W3 = OFF_B2 + H2;          // W3[OUT][H2]
constexpr int OFF_B3 = OFF_W3 + OUT * H2;    // b3[OUT]
constexpr int NPARAM = OFF_B3 + OUT;
static_assert(NPARAM == ML_MLP_NUM_PARAMS, "parameter count");
static_assert(H1 == 64 && H2 == 64, "one hidden unit per lane of a wave64");

constexpr int MLP_WAVES = 4;  // waves per workgroup

struct MlpLds {
  float w1t[IN][H1];   // W1 transposed: w1t[i][j] = W1[j][i]
  float w2[H2][H1];    // W2 as stored: w2[j][k]
  float w2t[H1][H2];   // W2 transposed: w2t[k][j] = W2[j][k]
  float w3[OUT][H2];
  float b1[H1], b2[H2], b3[OUT];
};

__device__ __forceinline__ void load_weights(const float* p, MlpLds& s) {
  for (int e = threadIdx.x; e < H1 * IN; e += blockDim.x) {
    const int j = e / IN, i = e % IN;
    s.w1t[i][j] = p[OFF_W1 + e];
  }
  for (int e = threadIdx.x; e < H2 * H1; e += blockDim.x) {
    const int j = e / H1, k = e % H1;
    const float w = p[OFF_W2 + e];
    s.w2[j][k] = w;
    s.w2t[k][j] = w;
  }
  for (int e = threadIdx.x; e < OUT * H2; e += blockDim.x) s.w3[e / H2][e % H2] = p[OFF_W3 + e];
  for (int e = threadIdx.x; e < H1; e += blockDim.x) s.b1[e] = p[OFF_B1 + e];
  for (int e = threadIdx.x; e < H2; e += blockDim.x) s.b2[e] = p[OFF_B2 + e];
  if (threadIdx.x < OUT) s.b3[threadIdx.x] = p[OFF_B3 + threadIdx.x];
  __syncthreads();
}

__device__ __forceinline__ float wave_sum(float v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// Forward for one sample held in x (uniform); lane = hidden unit.
__device__ __forceinline__ void forward_one(const MlpLds& s, const float* x, int lane, float& h1, float& h2, float& y0,
                                            float& y1) {
  float a = s.b1[lane];
#pragma unroll
  for (int i = 0; i < IN; i++) a = fmaf(s.w1t[i][lane], x[i], a);
  h1 = a > 0.f ? a : 0.f;
  float b = s.b2[lane];
#pragma unroll
  for (int k = 0; k < H1; k++) b = fmaf(s.w2t[k][lane], __shfl(h1, k), b);
  h2 = b > 0.f ? b : 0.f;
  y0 = wave_sum(s.w3[0][lane] * h2) + s.b3[0];
  y1 = wave_sum(s.w3[1][lane] * h2) + s.b3[1];
}

__global__ void __launch_bounds__(64 * MLP_WAVES) mlp_forward_kernel(const float* params, const float* x, int64_t n,
                                                                     float* y) {
  __shared__ MlpLds s;
  load_weights(params, s);
  const int lane = threadIdx.x & 63;
  const int64_t wave = (int64_t)blockIdx.x * MLP_WAVES + (threadIdx.x >> 6);
  const int64_t nw = (int64_t)gridDim.x * MLP_WAVES;
  for (int64_t r = wave; r < n; r += nw) {
    float h1, h2, y0, y1;
    forward_one(s, x + r * IN, lane, h1, h2, y0, y1);
    if (lane == 0) {
      y[r * OUT + 0] = y0;
      y[r * OUT + 1] = y1;
    }
  }
}

// Loss = mean over samples and outputs of (y - t)^2. Per-wave partial
// gradients go to partial[wave][NPARAM]; partial_loss[wave] the squared error sum.
__global__ void __launch_bounds__(64 * MLP_WAVES) mlp_backward_kernel(const float* params, const float* x,
                                                                      const float* t, int64_t n, float* partial,
                                                                      float* partial_loss) {
  __shared__ MlpLds s;
  load_weights(params, s);
  const int lane = threadIdx.x & 63;
  const int64_t wave = (int64_t)blockIdx.x * MLP_WAVES + (threadIdx.x >> 6);
  const int64_t nw = (int64_t)gridDim.x * MLP_WAVES;
  const float scale = 2.0f / (float)(n * OUT);
  float gw1[IN], gw2[H1];
#pragma unroll
  for (int i = 0; i < IN; i++) gw1[i] = 0.f;
#pragma unroll
  for (int k = 0; k < H1; k++) gw2[k] = 0.f;
  float gb1 = 0.f, gb2 = 0.f, gw3_0 = 0.f, gw3_1 = 0.f, gb3_0 = 0.f, gb3_1 = 0.f, loss = 0.f;
  for (int64_t r = wave; r < n; r += nw) {
    const float* xr = x + r * IN;
    float h1, h2, y0, y1;
    forward_one(s, xr, lane, h1, h2, y0, y1);
    const float e0 = y0 - t[r * OUT + 0], e1 = y1 - t[r * OUT + 1];
    loss += e0 * e0 + e1 * e1;
    const float d0 = scale * e0, d1 = scale * e1;
    gb3_0 += d0;
    gb3_1 += d1;
    gw3_0 = fmaf(d0, h2, gw3_0);
    gw3_1 = fmaf(d1, h2, gw3_1);
    const float dh2 = h2 > 0.f ? fmaf(s.w3[0][lane], d0, s.w3[1][lane] * d1) : 0.f;
    gb2 += dh2;
#pragma unroll
    for (int k = 0; k < H1; k++) gw2[k] = fmaf(dh2, __shfl(h1, k), gw2[k]);
    float dh1 = 0.f;
#pragma unroll
    for (int j = 0; j < H2; j++) dh1 = fmaf(s.w2[j][lane], __shfl(dh2, j), dh1);
    dh1 = h1 > 0.f ? dh1 : 0.f;
    gb1 += dh1;
#pragma unroll
    for (int i = 0; i < IN; i++) gw1[i] = fmaf(dh1, xr[i], gw1[i]);
  }
  float* g = partial + wave * NPARAM;
#pragma unroll
  for (int i = 0; i < IN; i++) g[OFF_W1 + lane * IN + i] = gw1[i];
  g[OFF_B1 + lane] = gb1;
#pragma unroll
  for (int k = 0; k < H1; k++) g[OFF_W2 + lane * H1 + k] = gw2[k];
  g[OFF_B2 + lane] = gb2;
  g[OFF_W3 + lane] = gw3_0;
  g[OFF_W3 + H2 + lane] = gw3_1;
  if (lane == 0) {
    g[OFF_B3 + 0] = gb3_0;
    g[OFF_B3 + 1] = gb3_1;
    partial_loss[wave] = loss;
  }
}

// grads[p] = sum over waves (fixed order); loss[0] = mean squared error.
__global__ void mlp_reduce_kernel(const float* partial, const float* partial_loss, int nwaves, int64_t n, float* grads,
                                  float* loss) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < NPARAM) {
    float acc = 0.f;
    for (int w = 0; w < nwaves; w++) acc += partial[(size_t)w * NPARAM + p];
    grads[p] = acc;
  }
  if (p == 0) {
    double l = 0;
    for (int w = 0; w < nwaves; w++) l += partial_loss[w];
    loss[0] = (float)(l / (double)(n * OUT));
  }
}

// Adam (Kingma & Ba, 2015) with bias correction.
__global__ void mlp_adam_kernel(float* params, const float* grads, float* m, float* v, float lr, float b1, float b2,
                                float eps, float c1, float c2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= NPARAM) return;
  const float g = grads[p];
  const float mm = b1 * m[p] + (1.f - b1) * g;
  const float vv = b2 * v[p] + (1.f - b2) * g * g;
  m[p] = mm;
  v[p] = vv;
  params[p] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

// ------------------------------------------------------------------- features
// Per row group: min, max of an int64 key column (stats.rs ColumnStats shape).
__global__ void rowgroup_minmax_kernel(const int64_t* keys, int64_t n, int64_t rg_rows, int64_t* mins, int64_t* maxs) {
  __shared__ int64_t smin[256], smax[256];
  const int64_t g = blockIdx.x;
  const int64_t b = g * rg_rows, e = (b + rg_rows < n) ? b + rg_rows : n;
  int64_t lo = INT64_MAX, hi = INT64_MIN;
  for (int64_t i = b + threadIdx.x; i < e; i += blockDim.x) {
    const int64_t k = keys[i];
    lo = k < lo ? k : lo;
    hi = k > hi ? k : hi;
  }
  smin[threadIdx.x] = lo;
  smax[threadIdx.x] = hi;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      smin[threadIdx.x] = smin[threadIdx.x + o] < smin[threadIdx.x] ? smin[threadIdx.x + o] : smin[threadIdx.x];
      smax[threadIdx.x] = smax[threadIdx.x + o] > smax[threadIdx.x] ? smax[threadIdx.x + o] : smax[threadIdx.x];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    mins[g] = smin[0];
    maxs[g] = smax[0];
  }
}

// Sample rows i_k = splitmix64(seed + k) mod n and return their keys.
__device__ __forceinline__ uint64_t splitmix(uint64_t x) {
  uint64_t z = x + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__global__ void sample_keys_kernel(const int64_t* keys, int64_t n, int s, uint64_t seed, int64_t* out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= s) return;
  const uint64_t r = splitmix(seed + (uint64_t)k);
  const int64_t i = (int64_t)(((unsigned __int128)r * (uint64_t)n) >> 64);
  out[k] = keys[i];
}

}  // namespace mlg

// ------------------------------------------------------------------- host side
namespace mlg {
struct MlpModel {
  ml_ctx* ctx = nullptr;
  Buffer params, grads, m, v, loss;
  int64_t step = 0;
};
}  // namespace mlg

struct ml_mlp {
  mlg::MlpModel model;
};

using namespace mlg;

// Features from row-group min/max and a key sample (all logs base 2):
//  0 rows, 1 row groups, 2 global key range, 3/4 mean/std of row-group ranges,
//  5 sum of row-group ranges / global range, 6 share of constant row groups,
//  7 min(range, rows), 8 sample size, 9 sample distinct d, 10 d/s, 11-13 f1,f2,f3 / s
//  (values seen once/twice/three times), 14 max frequency / s, 15 GEE estimate
//  (Charikar et al. 2000), 16 Chao1 (Chao 1984), 17 d * rows / s, 18 Shlosser
//  (1981), 19-23 KMV estimates from the k-th smallest distinct hash,
//  k = 1,4,16,64,256 (Bar-Yossef et al. 2002), 24 range / rows, 25 normalized
//  sample entropy, 26 share of values seen >= 4 times, 27-31 left for the caller
//  (join-side features).
static void sizing_features_host(int64_t n, int64_t ngroups, const int64_t* mins, const int64_t* maxs,
                                 const int64_t* sample, int s, float* f) {
  auto lg = [](double x) { return (float)std::log2(x < 1.0 ? 1.0 : x); };
  int64_t gmin = INT64_MAX, gmax = INT64_MIN;
  double sum_r = 0, sum_l = 0, sum_l2 = 0;
  int64_t constant = 0;
  for (int64_t g = 0; g < ngroups; g++) {
    gmin = std::min(gmin, mins[g]);
    gmax = std::max(gmax, maxs[g]);
    const double r = (double)maxs[g] - (double)mins[g] + 1.0;
    sum_r += r;
    const double l = std::log2(r);
    sum_l += l;
    sum_l2 += l * l;
    constant += maxs[g] == mins[g];
  }
  const double R = (double)gmax - (double)gmin + 1.0;
  const double ml = sum_l / (double)ngroups;
  f[0] = lg((double)n);
  f[1] = lg((double)ngroups);
  f[2] = lg(R);
  f[3] = (float)ml;
  f[4] = (float)std::sqrt(std::max(0.0, sum_l2 / (double)ngroups - ml * ml));
  f[5] = lg(sum_r / R);
  f[6] = (float)((double)constant / (double)ngroups);
  f[7] = lg(std::min(R, (double)n));
  std::vector<int64_t> v(sample, sample + s);
  std::sort(v.begin(), v.end());
  std::vector<int64_t> freq;
  for (size_t i = 0; i < v.size();) {
    size_t j = i;
    while (j < v.size() && v[j] == v[i]) j++;
    freq.push_back((int64_t)(j - i));
    i = j;
  }
  const double d = (double)freq.size();
  double f1 = 0, f2 = 0, f3 = 0, f4p = 0, mx = 0, ent = 0;
  for (int64_t c : freq) {
    f1 += c == 1;
    f2 += c == 2;
    f3 += c == 3;
    f4p += c >= 4;
    mx = std::max(mx, (double)c);
    const double p = (double)c / s;
    ent -= p * std::log2(p);
  }
  const double q = std::min(1.0, (double)s / (double)n);
  f[8] = lg(s);
  f[9] = lg(d);
  f[10] = (float)(d / s);
  f[11] = (float)(f1 / s);
  f[12] = (float)(f2 / s);
  f[13] = (float)(f3 / s);
  f[14] = (float)(mx / s);
  f[15] = lg(std::sqrt(1.0 / q) * f1 + (d - f1));
  f[16] = lg(d + f1 * (f1 - 1) / (2 * (f2 + 1)));
  f[17] = lg(std::min(d / q, (double)n));
  {
    double num = 0, den = 0;
    std::vector<double> fi(mx > 0 ? (size_t)mx + 1 : 1, 0);
    for (int64_t c : freq) fi[(size_t)c] += 1;
    for (size_t i = 1; i < fi.size(); i++) {
      num += std::pow(1 - q, (double)i) * fi[i];
      den += (double)i * q * std::pow(1 - q, (double)i - 1) * fi[i];
    }
    f[18] = lg(d + (den > 0 ? f1 * num / den : 0));
  }
  std::vector<uint64_t> hs;
  for (size_t i = 0; i < v.size(); i++)
    if (i == 0 || v[i] != v[i - 1]) hs.push_back(hash_u64((uint64_t)v[i]));
  std::sort(hs.begin(), hs.end());
  const int ks[5] = {1, 4, 16, 64, 256};
  for (int t = 0; t < 5; t++) {
    const size_t k = (size_t)ks[t];
    if (hs.size() >= k) {
      const double u = ((double)hs[k - 1] + 1.0) / 18446744073709551616.0;
      f[19 + t] = lg((double)k / u);
    } else {
      f[19 + t] = lg((double)hs.size());
    }
  }
  f[24] = (float)std::log2(1.0 + R / (double)n);
  f[25] = (float)(s > 1 ? ent / std::log2((double)s) : 0.0);
  f[26] = (float)(f4p / s);
}

template <class F>
static ml_status mlp_guard(ml_ctx* ctx, F f) {
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

static int mlp_grid(ml_ctx* ctx, int64_t n) {
  int64_t waves = std::min<int64_t>(n, (int64_t)ctx->num_cus * 2 * MLP_WAVES);
  int64_t g = (waves + MLP_WAVES - 1) / MLP_WAVES;
  return (int)std::max<int64_t>(1, g);
}

static void mlp_grads(MlpModel& M, const float* x, const float* t, int64_t n) {
  ml_ctx* ctx = M.ctx;
  const int grid = mlp_grid(ctx, n);
  const int nwaves = grid * MLP_WAVES;
  Buffer partial(ctx->arena, (size_t)nwaves * NPARAM * 4);
  Buffer ploss(ctx->arena, (size_t)nwaves * 4);
  mlp_backward_kernel<<<grid, 64 * MLP_WAVES, 0, ctx->stream>>>(M.params.as<float>(), x, t, n, partial.as<float>(),
                                                               ploss.as<float>());
  mlp_reduce_kernel<<<(NPARAM + 255) / 256, 256, 0, ctx->stream>>>(partial.as<float>(), ploss.as<float>(), nwaves, n,
                                                                   M.grads.as<float>(), M.loss.as<float>());
}

extern "C" {

ml_status ml_mlp_create(ml_ctx* ctx, const float* params, ml_mlp** out) {
  auto* m = new ml_mlp();
  ml_status st = mlp_guard(ctx, [&] {
    MlpModel& M = m->model;
    M.ctx = ctx;
    M.params = Buffer(ctx->arena, NPARAM * 4);
    M.grads = Buffer(ctx->arena, NPARAM * 4);
    M.m = Buffer(ctx->arena, NPARAM * 4);
    M.v = Buffer(ctx->arena, NPARAM * 4);
    M.loss = Buffer(ctx->arena, 256);
    hip_check(hipMemcpyAsync(M.params.ptr, params, NPARAM * 4, hipMemcpyHostToDevice, ctx->stream), "upload");
    launch_fill_u64(M.m.as<uint64_t>(), 0, (NPARAM * 4 + 7) / 8, ctx->stream);
    launch_fill_u64(M.v.as<uint64_t>(), 0, (NPARAM * 4 + 7) / 8, ctx->stream);
    hip_check(hipStreamSynchronize(ctx->stream), "sync");
  });
  if (st != ML_OK) {
    delete m;
    *out = nullptr;
    return st;
  }
  *out = m;
  return ML_OK;
}

ml_status ml_mlp_get_params(ml_mlp* m, float* params) {
  ml_ctx* ctx = m->model.ctx;
  return mlp_guard(ctx, [&] {
    hip_check(hipMemcpyAsync(params, m->model.params.ptr, NPARAM * 4, hipMemcpyDeviceToHost, ctx->stream), "download");
    hip_check(hipStreamSynchronize(ctx->stream), "sync");
  });
}

ml_status ml_mlp_forward(ml_mlp* m, const float* x, int64_t n, float* y) {
  ml_ctx* ctx = m->model.ctx;
  return mlp_guard(ctx, [&] {
    if (n <= 0) return;
    mlp_forward_kernel<<<mlp_grid(ctx, n), 64 * MLP_WAVES, 0, ctx->stream>>>(m->model.params.as<float>(), x, n, y);
    hip_check(hipGetLastError(), "mlp_forward launch");
  });
}

ml_status ml_mlp_loss_grad(ml_mlp* m, const float* x, const float* t, int64_t n, float* loss, float* grads_host) {
  ml_ctx* ctx = m->model.ctx;
  return mlp_guard(ctx, [&] {
    if (n <= 0) fail(ML_ERR_EXECUTION, "empty batch");
    mlp_grads(m->model, x, t, n);
    hip_check(hipMemcpyAsync(grads_host, m->model.grads.ptr, NPARAM * 4, hipMemcpyDeviceToHost, ctx->stream), "copy");
    hip_check(hipMemcpyAsync(loss, m->model.loss.ptr, 4, hipMemcpyDeviceToHost, ctx->stream), "copy");
    hip_check(hipStreamSynchronize(ctx->stream), "sync");
  });
}

ml_status ml_mlp_train_step(ml_mlp* m, const float* x, const float* t, int64_t n, float lr, float* loss) {
  ml_ctx* ctx = m->model.ctx;
  return mlp_guard(ctx, [&] {
    if (n <= 0) fail(ML_ERR_EXECUTION, "empty batch");
    MlpModel& M = m->model;
    mlp_grads(M, x, t, n);
    M.step++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - (float)std::pow((double)b1, (double)M.step);
    const float c2 = 1.f - (float)std::pow((double)b2, (double)M.step);
    mlp_adam_kernel<<<(NPARAM + 255) / 256, 256, 0, ctx->stream>>>(M.params.as<float>(), M.grads.as<float>(),
                                                                   M.m.as<float>(), M.v.as<float>(), lr, b1, b2, 1e-8f,
                                                                   c1, c2);
    if (loss) {
      hip_check(hipMemcpyAsync(loss, M.loss.ptr, 4, hipMemcpyDeviceToHost, ctx->stream), "copy");
      hip_check(hipStreamSynchronize(ctx->stream), "sync");
    }
  });
}

void ml_mlp_free(ml_mlp* m) { delete m; }

// 32 features of an int64 key column: row-group statistics (min/max per row
// group, the ColumnStats of crates/minilake-storage/src/stats.rs:15-22) and a
// `sample`-row sample with its KMV sketch and frequency profile.
ml_status ml_sizing_features(ml_ctx* ctx, const int64_t* keys, int64_t n, int64_t row_group_rows, int32_t sample,
                             uint64_t seed, float* out) {
  return mlp_guard(ctx, [&] {
    for (int i = 0; i < ML_MLP_IN; i++) out[i] = 0.f;
    if (n <= 0) return;
    if (row_group_rows <= 0) row_group_rows = 122880;
    if (sample <= 0) sample = 4096;
    const int64_t ngroups = (n + row_group_rows - 1) / row_group_rows;
    Buffer mins(ctx->arena, (size_t)ngroups * 8), maxs(ctx->arena, (size_t)ngroups * 8);
    rowgroup_minmax_kernel<<<(unsigned)ngroups, 256, 0, ctx->stream>>>(keys, n, row_group_rows, mins.as<int64_t>(),
                                                                       maxs.as<int64_t>());
    Buffer samp(ctx->arena, (size_t)sample * 8);
    sample_keys_kernel<<<(sample + 255) / 256, 256, 0, ctx->stream>>>(keys, n, sample, seed, samp.as<int64_t>());
    std::vector<int64_t> hmin((size_t)ngroups), hmax((size_t)ngroups), hs((size_t)sample);
    hip_check(hipMemcpyAsync(hmin.data(), mins.ptr, hmin.size() * 8, hipMemcpyDeviceToHost, ctx->stream), "copy");
    hip_check(hipMemcpyAsync(hmax.data(), maxs.ptr, hmax.size() * 8, hipMemcpyDeviceToHost, ctx->stream), "copy");
    hip_check(hipMemcpyAsync(hs.data(), samp.ptr, hs.size() * 8, hipMemcpyDeviceToHost, ctx->stream), "copy");
    hip_check(hipStreamSynchronize(ctx->stream), "sync");
    sizing_features_host(n, ngroups, hmin.data(), hmax.data(), hs.data(), sample, out);
  });
}

ml_status ml_ctx_set_estimator(ml_ctx* ctx, ml_mlp* m) {
  return mlp_guard(ctx, [&] { ctx->estimator = m ? &m->model : nullptr; });
}

}  // extern "C"
