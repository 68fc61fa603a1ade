// CPU ORACLE — test infrastructure only (see ref_core.h).
// Folded-multiply hash and SwissTable, restating
// crates/minilake-hashtable/src/hash.rs:12-55 and swiss.rs:34-301.
#pragma once

#include <cstdint>
#include <cstring>
#include <vector>

namespace ref {

// hash.rs:12-17
constexpr uint64_t SEED = 0x243F6A8885A308D3ull;
constexpr uint64_t MUL = 0x9E3779B97F4A7C15ull;
constexpr uint64_t MUL2 = 0xBF58476D1CE4E5B9ull;
constexpr uint64_t NULL_HASH = 0x7C159E3779B97F4Aull;

// hash.rs:21-24: low 64 bits XOR high 64 bits of the 128-bit product.
inline uint64_t fold_mul(uint64_t a, uint64_t b) {
  unsigned __int128 r = (unsigned __int128)a * (unsigned __int128)b;
  return (uint64_t)r ^ (uint64_t)(r >> 64);
}
// hash.rs:28-30
inline uint64_t hash_u64(uint64_t x) { return fold_mul(x ^ SEED, MUL); }
// hash.rs:34-36
inline uint64_t combine(uint64_t h, uint64_t v) { return fold_mul(h ^ v, MUL2); }
// hash.rs:39-55: 8-byte little-endian words with MUL, zero-padded tail with MUL2.
inline uint64_t hash_bytes(const uint8_t* b, size_t n) {
  uint64_t h = SEED ^ ((uint64_t)n * MUL);
  size_t full = n / 8;
  for (size_t c = 0; c < full; c++) {
    uint64_t w;
    std::memcpy(&w, b + 8 * c, 8);
    h = fold_mul(h ^ w, MUL);
  }
  size_t tail = n - 8 * full;
  if (tail) {
    uint64_t w = 0;
    std::memcpy(&w, b + 8 * full, tail);
    h = fold_mul(h ^ w, MUL2);
  }
  return h;
}

// ---------------------------------------------------------------------
// SwissTable (swiss.rs:50-56): ctrl bytes (0x80 = EMPTY, else 7-bit tag),
// full hashes and u32 payloads; power-of-two capacity in groups of 8.
// ---------------------------------------------------------------------
constexpr size_t GROUP = 8;
constexpr uint8_t EMPTY = 0x80;
constexpr uint64_t LSB = 0x0101010101010101ull;
constexpr uint64_t MSB = 0x8080808080808080ull;

inline uint8_t h2(uint64_t hash) { return (uint8_t)(hash >> 57); }  // swiss.rs:59-61
// swiss.rs:67-70 (may report a false positive just above a true match).
inline uint64_t match_tag(uint64_t group, uint8_t tag) {
  uint64_t x = group ^ (LSB * (uint64_t)tag);
  return (x - LSB) & ~x & MSB;
}
inline uint64_t match_empty(uint64_t group) { return group & MSB; }  // swiss.rs:74-76

inline size_t next_pow2(size_t x) {
  size_t p = 1;
  while (p < x) p <<= 1;
  return p;
}

struct SwissTable {
  std::vector<uint8_t> ctrl;
  std::vector<uint64_t> hashes;
  std::vector<uint32_t> payloads;
  size_t len = 0;
  size_t group_mask = 0;
  int resizes = 0;  // grow-and-rehash steps (instrumentation only)

  // swiss.rs:86-96: room for n entries at load <= 7/8, at least 2 groups.
  static SwissTable with_capacity(size_t n) {
    SwissTable t;
    size_t slots = next_pow2(n * 8 / 7 + 1);
    if (slots < GROUP * 2) slots = GROUP * 2;
    t.ctrl.assign(slots, EMPTY);
    t.hashes.assign(slots, 0);
    t.payloads.assign(slots, 0);
    t.group_mask = slots / GROUP - 1;
    return t;
  }
  size_t capacity() const { return ctrl.size(); }
  size_t memory_size() const { return capacity() * 13; }  // swiss.rs:114-116

  uint64_t load_group(size_t g) const {  // swiss.rs:129-134
    uint64_t w;
    std::memcpy(&w, &ctrl[g * GROUP], 8);
    return w;
  }

  // swiss.rs:138-158. Returns true and sets *out when found.
  template <class Eq>
  bool find(uint64_t hash, Eq eq, uint32_t* out) const {
    uint8_t tag = h2(hash);
    size_t g = (size_t)hash & group_mask;
    size_t stride = 0;
    for (;;) {
      uint64_t group = load_group(g);
      uint64_t m = match_tag(group, tag);
      while (m) {
        size_t i = g * GROUP + (size_t)(__builtin_ctzll(m) / 8);
        if (ctrl[i] == tag && hashes[i] == hash && eq(payloads[i])) {
          *out = payloads[i];
          return true;
        }
        m &= m - 1;
      }
      if (match_empty(group)) return false;
      stride += 1;
      g = (g + stride) & group_mask;
    }
  }

  // swiss.rs:163-185. Returns true (found, *payload set) or false (*slot = vacant slot).
  template <class Eq>
  bool find_or_vacant(uint64_t hash, Eq eq, uint32_t* payload, size_t* slot) {
    reserve(1);
    uint8_t tag = h2(hash);
    size_t g = (size_t)hash & group_mask;
    size_t stride = 0;
    for (;;) {
      uint64_t group = load_group(g);
      uint64_t m = match_tag(group, tag);
      while (m) {
        size_t i = g * GROUP + (size_t)(__builtin_ctzll(m) / 8);
        if (ctrl[i] == tag && hashes[i] == hash && eq(payloads[i])) {
          *payload = payloads[i];
          return true;
        }
        m &= m - 1;
      }
      uint64_t empty = match_empty(group);
      if (empty) {
        *slot = g * GROUP + (size_t)(__builtin_ctzll(empty) / 8);
        return false;
      }
      stride += 1;
      g = (g + stride) & group_mask;
    }
  }

  void insert_at(size_t slot, uint64_t hash, uint32_t payload) {  // swiss.rs:190-196
    ctrl[slot] = h2(hash);
    hashes[slot] = hash;
    payloads[slot] = payload;
    len += 1;
  }

  size_t first_empty(uint64_t hash) const {  // swiss.rs:205-216
    size_t g = (size_t)hash & group_mask;
    size_t stride = 0;
    for (;;) {
      uint64_t empty = match_empty(load_group(g));
      if (empty) return g * GROUP + (size_t)(__builtin_ctzll(empty) / 8);
      stride += 1;
      g = (g + stride) & group_mask;
    }
  }

  // swiss.rs:219-232: double (at least) and rehash when len+extra exceeds 7/8.
  void reserve(size_t extra) {
    size_t needed = len + extra;
    if (needed * 8 <= capacity() * 7) return;
    SwissTable bigger = with_capacity(needed > len * 2 ? needed : len * 2);
    for (size_t i = 0; i < capacity(); i++) {
      if (ctrl[i] != EMPTY) {
        size_t s = bigger.first_empty(hashes[i]);
        bigger.insert_at(s, hashes[i], payloads[i]);
      }
    }
    bigger.resizes = resizes + 1;
    *this = std::move(bigger);
  }

  // swiss.rs:240-281: KeyStore provides equals(payload, k) and insert(k).
  template <class KeyStore>
  size_t find_or_insert_batch(const uint64_t* hs, size_t n, KeyStore& keys, std::vector<uint32_t>& out) {
    reserve(n);
    out.clear();
    out.reserve(n);
    size_t inserted = 0;
    for (size_t k = 0; k < n; k++) {
      uint64_t h = hs[k];
      uint8_t tag = h2(h);
      size_t g = (size_t)h & group_mask;
      size_t stride = 0;
      uint32_t payload = 0;
      for (;;) {
        uint64_t group = load_group(g);
        uint64_t m = match_tag(group, tag);
        bool found = false;
        while (m) {
          size_t i = g * GROUP + (size_t)(__builtin_ctzll(m) / 8);
          if (ctrl[i] == tag && hashes[i] == h && keys.equals(payloads[i], k)) {
            payload = payloads[i];
            found = true;
            break;
          }
          m &= m - 1;
        }
        if (found) break;
        uint64_t empty = match_empty(group);
        if (empty) {
          size_t slot = g * GROUP + (size_t)(__builtin_ctzll(empty) / 8);
          payload = keys.insert(k);
          insert_at(slot, h, payload);
          inserted++;
          break;
        }
        stride += 1;
        g = (g + stride) & group_mask;
      }
      out.push_back(payload);
    }
    return inserted;
  }
};

}  // namespace ref
