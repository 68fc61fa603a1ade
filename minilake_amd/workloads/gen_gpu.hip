// Device-side generator for the large synthetic key columns of the
// learned-sizing training set and evaluation workload. The same counter-based
// splitmix64 streams, Zipf rejection-inversion (Hörmann & Derflinger 1996) and
// Feistel-network key scramble as gen.cpp, so the host and device generators
// produce identical columns for the same (seed, stream, row).
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

__host__ __device__ inline uint64_t splitmix64(uint64_t x) {
  uint64_t z = x + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__host__ __device__ inline uint64_t stream_key(uint64_t seed, uint64_t stream) {
  return splitmix64(splitmix64(seed) ^ (stream * 0xD1B54A32D192ED03ull));
}
__host__ __device__ inline uint64_t draw(uint64_t key, uint64_t i) { return splitmix64(key + i * 0x9E3779B97F4A7C15ull); }
__device__ inline uint64_t below(uint64_t r, uint64_t n) { return __umul64hi(r, n); }
__device__ inline double unit(uint64_t r) { return (double)(r >> 11) * (1.0 / 9007199254740992.0); }

struct Zipf {
  double s, n, hx1, hn, sq;
};
__host__ __device__ inline double helper1(double x) {
  return fabs(x) > 1e-8 ? log1p(x) / x : 1 - x * (0.5 - x * (1.0 / 3 - 0.25 * x));
}
__host__ __device__ inline double helper2(double x) {
  return fabs(x) > 1e-8 ? expm1(x) / x : 1 + x * 0.5 * (1 + x * (1.0 / 3) * (1 + 0.25 * x));
}
__host__ __device__ inline double zh(const Zipf& z, double x) { return exp(-z.s * log(x)); }
__host__ __device__ inline double zH(const Zipf& z, double x) {
  const double lx = log(x);
  return helper2((1 - z.s) * lx) * lx;
}
__host__ __device__ inline double zHinv(const Zipf& z, double x) {
  double t = x * (1 - z.s);
  if (t < -1) t = -1;
  return exp(helper1(t) * x);
}
Zipf make_zipf(double s, double n) {
  Zipf z{s, n, 0, 0, 0};
  z.hx1 = zH(z, 1.5) - 1;
  z.hn = zH(z, n + 0.5);
  z.sq = 2 - zHinv(z, zH(z, 2.5) - zh(z, 2));
  return z;
}
__device__ inline uint64_t zipf_sample(const Zipf& z, uint64_t key, uint64_t i) {
  for (uint64_t attempt = 0;; attempt++) {
    const double u = z.hn + unit(draw(key, i * 64 + attempt)) * (z.hx1 - z.hn);
    const double x = zHinv(z, u);
    double k = floor(x + 0.5);
    if (k < 1) k = 1;
    if (k > z.n) k = z.n;
    if (k - x <= z.sq || u >= zH(z, k + 0.5) - zh(z, k)) return (uint64_t)k;
    if (attempt == 63) return (uint64_t)k;
  }
}

// Feistel permutation of {0..n-1} with cycle walking (matches gen.cpp mlw_permute).
__device__ inline uint64_t permute(uint64_t key, int half, uint64_t mask, uint64_t n, uint64_t x) {
  do {
    uint64_t l = x >> half, r = x & mask;
    for (int round = 0; round < 4; round++) {
      const uint64_t f = draw(key, (r << 2) | (uint64_t)round) & mask;
      const uint64_t nl = r, nr = l ^ f;
      l = nl;
      r = nr;
    }
    x = (l << half) | r;
  } while (x >= n);
  return x;
}

__global__ void zipf_keys_kernel(int64_t* out, int64_t n, uint64_t zkey, Zipf z, int uniform, uint64_t pkey, int half,
                                 uint64_t pmask, uint64_t domain) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint64_t rank = uniform ? 1 + below(draw(zkey, (uint64_t)i), domain) : zipf_sample(z, zkey, (uint64_t)i);
    out[i] = (int64_t)permute(pkey, half, pmask, domain, rank - 1);
  }
}

__global__ void build_keys_kernel(int64_t* out, int64_t nkeys, int64_t mult, uint64_t pkey, int half, uint64_t pmask,
                                  uint64_t domain) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < nkeys * mult; j += stride)
    out[j] = (int64_t)permute(pkey, half, pmask, domain, (uint64_t)(j / mult));
}

// One batched-fragment workload fragment: build keys unique (the Feistel
// permutation of the build row over [0, 2^62)), build groups uniform in
// [0, groups), probe keys the build key of a Zipf(s)-distributed build row
// (s <= 0: uniform), probe values uniform in [0, 1).
__global__ void fragment_kernel(int64_t* bkeys, int64_t* bgroups, int64_t brows, int64_t* pkeys, double* pvals,
                                int64_t prows, uint64_t groups, uint64_t kkey, int khalf, uint64_t kmask,
                                uint64_t kdomain, uint64_t gkey, uint64_t zkey, Zipf z, int uniform, uint64_t vkey) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < brows; j += stride) {
    bkeys[j] = (int64_t)permute(kkey, khalf, kmask, kdomain, (uint64_t)j);
    bgroups[j] = (int64_t)below(draw(gkey, (uint64_t)j), groups);
  }
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < prows; i += stride) {
    const uint64_t rank = uniform ? 1 + below(draw(zkey, (uint64_t)i), (uint64_t)brows) : zipf_sample(z, zkey, (uint64_t)i);
    pkeys[i] = (int64_t)permute(kkey, khalf, kmask, kdomain, rank - 1);
    pvals[i] = unit(draw(vkey, (uint64_t)i));
  }
}

void feistel_params(uint64_t n, int* half, uint64_t* mask) {
  int bits = 2;
  while ((1ull << bits) < n) bits += 2;
  *half = bits / 2;
  *mask = (1ull << *half) - 1;
}

}  // namespace

extern "C" {

// out[i] = scramble(rank_i - 1), rank_i ~ Zipf(s) over {1..domain} (s <= 0: uniform),
// scramble = the Feistel permutation of stream (seed, perm_stream) over {0..domain-1}.
int mlw_gpu_zipf_keys(int64_t* out, int64_t n, uint64_t seed, uint64_t stream, double s, int64_t domain,
                      uint64_t perm_stream, void* stream_ptr) {
  if (n <= 0) return 0;
  if (domain <= 0) return 1;
  int half;
  uint64_t mask;
  feistel_params((uint64_t)domain, &half, &mask);
  const Zipf z = make_zipf(s > 0 ? s : 1.0, (double)domain);
  const int64_t blocks = (n + 255) / 256 < 65536 ? (n + 255) / 256 : 65536;
  zipf_keys_kernel<<<(unsigned)blocks, 256, 0, (hipStream_t)stream_ptr>>>(
      out, n, stream_key(seed, stream), z, s <= 0 ? 1 : 0, stream_key(seed, perm_stream), half, mask, (uint64_t)domain);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

// out[j*mult + r] = scramble(j) for j < nkeys (each key `mult` times). The
// keys are distinct values of {0..domain-1}, so nkeys may not exceed domain
// (the permutation is only defined on {0..domain-1}: a larger input would
// never leave its cycle walk); such a request returns 1 without a launch.
int mlw_gpu_build_keys(int64_t* out, int64_t nkeys, int64_t mult, uint64_t seed, uint64_t perm_stream, int64_t domain,
                       void* stream_ptr) {
  if (nkeys <= 0 || mult <= 0) return 0;
  if (domain <= 0 || nkeys > domain) return 1;
  int half;
  uint64_t mask;
  feistel_params((uint64_t)domain, &half, &mask);
  const int64_t n = nkeys * mult;
  const int64_t blocks = (n + 255) / 256 < 65536 ? (n + 255) / 256 : 65536;
  build_keys_kernel<<<(unsigned)blocks, 256, 0, (hipStream_t)stream_ptr>>>(out, nkeys, mult,
                                                                            stream_key(seed, perm_stream), half, mask,
                                                                            (uint64_t)domain);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

// Fragment `index` of the batched-fragment workload (see fragment_kernel);
// every column is a function of (seed, index, row) only.
int mlw_gpu_fragment(int64_t* bkeys, int64_t* bgroups, int64_t brows, int64_t* pkeys, double* pvals, int64_t prows,
                     int64_t groups, double s, uint64_t seed, uint64_t index, void* stream_ptr) {
  if (brows <= 0 || prows < 0 || groups <= 0) return 1;
  const uint64_t kdomain = 1ull << 62;
  int half;
  uint64_t mask;
  feistel_params(kdomain, &half, &mask);
  const Zipf z = make_zipf(s > 0 ? s : 1.0, (double)brows);
  const uint64_t base = 0x4652414700000000ull + index * 8;
  const int64_t n = brows > prows ? brows : prows;
  const int64_t blocks = (n + 255) / 256 < 16384 ? (n + 255) / 256 : 16384;
  fragment_kernel<<<(unsigned)(blocks > 0 ? blocks : 1), 256, 0, (hipStream_t)stream_ptr>>>(
      bkeys, bgroups, brows, pkeys, pvals, prows, (uint64_t)groups, stream_key(seed, base), half, mask, kdomain,
      stream_key(seed, base + 1), stream_key(seed, base + 2), z, s <= 0 ? 1 : 0, stream_key(seed, base + 3));
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

}  // extern "C"
