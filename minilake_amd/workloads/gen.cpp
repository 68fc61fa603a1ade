// Seeded synthetic data for the benchmark and test workloads.
//
// Every value is a pure function of (seed, stream, row): a counter-based
// splitmix64 (Steele, Lea & Flood, "Fast splittable pseudorandom number
// generators", OOPSLA 2014), so any table can be regenerated exactly, in any
// order and on any number of threads. Zipf samples use Hörmann & Derflinger's
// rejection-inversion method ("Rejection-inversion to generate variates from
// monotone discrete distributions", ACM TOMACS 6(3), 1996).
//
// Host-only C++ (g++); exported with a C ABI for ctypes.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

namespace {

inline uint64_t splitmix64(uint64_t x) {
  uint64_t z = x + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
inline uint64_t stream_key(uint64_t seed, uint64_t stream) { return splitmix64(splitmix64(seed) ^ (stream * 0xD1B54A32D192ED03ull)); }
inline uint64_t draw(uint64_t key, uint64_t i) { return splitmix64(key + i * 0x9E3779B97F4A7C15ull); }
// uniform integer in [0, n) (multiply-shift; bias < n / 2^64)
inline uint64_t below(uint64_t r, uint64_t n) { return (uint64_t)(((unsigned __int128)r * n) >> 64); }
// uniform double in [0, 1) with 53 random bits
inline double unit(uint64_t r) { return (double)(r >> 11) * (1.0 / 9007199254740992.0); }

template <class F>
void parallel_for(int64_t n, F f) {
  int64_t threads = (int64_t)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (n < 200000) threads = 1;
  int64_t chunk = (n + threads - 1) / threads;
  std::vector<std::thread> ts;
  for (int64_t t = 0; t < threads; t++) {
    int64_t b = t * chunk, e = std::min(n, b + chunk);
    if (b >= e) break;
    ts.emplace_back([=] { for (int64_t i = b; i < e; i++) f(i); });
  }
  for (auto& t : ts) t.join();
}

// Zipf over {1..n} with exponent s > 0 (rejection-inversion).
struct Zipf {
  double s, n, hx1, hn, sq;
  static double helper1(double x) { return std::fabs(x) > 1e-8 ? std::log1p(x) / x : 1 - x * (0.5 - x * (1.0 / 3 - 0.25 * x)); }
  static double helper2(double x) { return std::fabs(x) > 1e-8 ? std::expm1(x) / x : 1 + x * 0.5 * (1 + x * (1.0 / 3) * (1 + 0.25 * x)); }
  double h(double x) const { return std::exp(-s * std::log(x)); }
  double H(double x) const { double lx = std::log(x); return helper2((1 - s) * lx) * lx; }
  double Hinv(double x) const { double t = x * (1 - s); if (t < -1) t = -1; return std::exp(helper1(t) * x); }
  Zipf(double exponent, double count) : s(exponent), n(count) {
    hx1 = H(1.5) - 1;
    hn = H(n + 0.5);
    sq = 2 - Hinv(H(2.5) - h(2));
  }
  uint64_t sample(uint64_t key, uint64_t i) const {
    for (uint64_t attempt = 0;; attempt++) {
      double u = hn + unit(draw(key, i * 64 + attempt)) * (hx1 - hn);
      double x = Hinv(u);
      double k = std::floor(x + 0.5);
      if (k < 1) k = 1;
      if (k > n) k = n;
      if (k - x <= sq || u >= H(k + 0.5) - h(k)) return (uint64_t)k;
      if (attempt == 63) return (uint64_t)k;  // vanishingly rare; keeps the loop bounded
    }
  }
};

}  // namespace

extern "C" {

// out[i] = lo + U{0 .. hi-lo} for rows offset .. offset+n-1 of the stream.
void mlw_uniform_int(uint64_t seed, uint64_t stream, int64_t offset, int64_t n, int64_t lo, int64_t hi, int64_t* out) {
  uint64_t key = stream_key(seed, stream);
  uint64_t span = (uint64_t)(hi - lo) + 1;
  parallel_for(n, [&](int64_t i) { out[i] = lo + (int64_t)below(draw(key, (uint64_t)(offset + i)), span); });
}

// out[i] = lo + u * (hi - lo), u uniform in [0, 1).
void mlw_uniform_f64(uint64_t seed, uint64_t stream, int64_t offset, int64_t n, double lo, double hi, double* out) {
  uint64_t key = stream_key(seed, stream);
  parallel_for(n, [&](int64_t i) { out[i] = lo + unit(draw(key, (uint64_t)(offset + i))) * (hi - lo); });
}

// out[i] = table[U{0 .. k-1}]
void mlw_pick_f64(uint64_t seed, uint64_t stream, int64_t offset, int64_t n, const double* table, int64_t k, double* out) {
  uint64_t key = stream_key(seed, stream);
  parallel_for(n, [&](int64_t i) { out[i] = table[below(draw(key, (uint64_t)(offset + i)), (uint64_t)k)]; });
}

// Zipf(s) over {1..domain}; s == 0 gives the uniform distribution.
void mlw_zipf(uint64_t seed, uint64_t stream, int64_t offset, int64_t n, double s, int64_t domain, int64_t* out) {
  uint64_t key = stream_key(seed, stream);
  if (s <= 0) {
    parallel_for(n, [&](int64_t i) { out[i] = 1 + (int64_t)below(draw(key, (uint64_t)(offset + i)), (uint64_t)domain); });
    return;
  }
  Zipf z(s, (double)domain);
  parallel_for(n, [&](int64_t i) { out[i] = (int64_t)z.sample(key, (uint64_t)(offset + i)); });
}

// Bijective scramble of {0..n-1} (a random-looking permutation): maps i to a
// distinct value using cycle-walking over a 4-round Feistel network on the
// next power of four.
void mlw_permute(uint64_t seed, uint64_t stream, int64_t offset, int64_t count, int64_t n, int64_t* out) {
  uint64_t key = stream_key(seed, stream);
  int bits = 2;
  while ((1ull << bits) < (uint64_t)n) bits += 2;
  int half = bits / 2;
  uint64_t mask = (1ull << half) - 1;
  parallel_for(count, [&](int64_t j) {
    uint64_t x = (uint64_t)(offset + j);
    do {
      uint64_t l = x >> half, r = x & mask;
      for (int round = 0; round < 4; round++) {
        uint64_t f = draw(key, (r << 2) | (uint64_t)round) & mask;
        uint64_t nl = r, nr = l ^ f;
        l = nl;
        r = nr;
      }
      x = (l << half) | r;
    } while (x >= (uint64_t)n);
    out[j] = (int64_t)x;
  });
}

// out[i] = the same Feistel permutation applied to values[i] (each in [0, n)).
void mlw_permute_values(uint64_t seed, uint64_t stream, int64_t count, int64_t n, const int64_t* values, int64_t* out) {
  uint64_t key = stream_key(seed, stream);
  int bits = 2;
  while ((1ull << bits) < (uint64_t)n) bits += 2;
  int half = bits / 2;
  uint64_t mask = (1ull << half) - 1;
  parallel_for(count, [&](int64_t j) {
    uint64_t x = (uint64_t)values[j];
    do {
      uint64_t l = x >> half, r = x & mask;
      for (int round = 0; round < 4; round++) {
        uint64_t f = draw(key, (r << 2) | (uint64_t)round) & mask;
        uint64_t nl = r, nr = l ^ f;
        l = nl;
        r = nr;
      }
      x = (l << half) | r;
    } while (x >= (uint64_t)n);
    out[j] = (int64_t)x;
  });
}

// Parallel gather: out[i] = src[idx[i]] for 8-byte and 4-byte elements.
void mlw_gather8(const int64_t* src, const int64_t* idx, int64_t n, int64_t* out) {
  parallel_for(n, [&](int64_t i) { out[i] = src[idx[i]]; });
}
void mlw_gather4(const int32_t* src, const int64_t* idx, int64_t n, int32_t* out) {
  parallel_for(n, [&](int64_t i) { out[i] = src[idx[i]]; });
}

}  // extern "C"
