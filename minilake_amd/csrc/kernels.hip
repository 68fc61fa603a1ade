// gfx950 kernels of the execution path: the fused pipeline interpreter and
// its sinks, hash-table build/rehash/finalize, ordered compaction and TopN.
//
// Layout and launch notes (MI355X, CDNA4):
//  * wave64; streaming kernels use 256-thread workgroups and grid-stride loops
//    over rows, sized to a multiple of the 256 CUs;
//  * expression programs are read with scalar loads (constant address space),
//    so instruction dispatch is uniform SALU work;
//  * FP64 accumulation uses hardware atomics (global_atomic_add_f64,
//    ds_add_f64); per-row arithmetic is compiled with -ffp-contract=off so
//    every row value is bit-identical to an IEEE evaluation in source order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <utility>

#include "kernels.h"
#include "mlgpu.h"

namespace mlg {

typedef __attribute__((address_space(4))) const Instr CInstr;
typedef __attribute__((address_space(4))) const uint64_t CU64;

// Device buffers are global memory. Pointers that reach a kernel through an
// instruction word or a void* field would otherwise be generic (flat)
// accesses, which also count against the scalar-load counter: every wait for
// an instruction fetch would then wait for all outstanding loads as well.
template <class T>
__device__ __forceinline__ const __attribute__((address_space(1))) T* gp(const void* p) {
  return (const __attribute__((address_space(1))) T*)p;
}
template <class T>
__device__ __forceinline__ __attribute__((address_space(1))) T* gpw(void* p) {
  return (__attribute__((address_space(1))) T*)p;
}

// A value every active lane holds; readfirstlane lets the compiler keep it
// (and whatever it indexes in constant memory) in scalar registers.
__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }

// The instruction stream of a plan. Instruction indices are the same in every
// lane; fetching through at() makes them scalar loads. The index is clamped to
// the program, so a fetch issued while no lane is active stays in the buffer.
struct Code {
  CInstr* base;
  int32_t last;
  __device__ __forceinline__ CInstr& at(int i) const {
    const int j = uni(i);
    return base[j < last ? j : last];
  }
};

__device__ __forceinline__ double u2f(uint64_t x) { return __longlong_as_double((long long)x); }
__device__ __forceinline__ uint64_t f2u(double x) { return (uint64_t)__double_as_longlong(x); }

// Row indices are 32-bit: sources and build sides are limited to 2^32 - 1 rows.
__device__ __forceinline__ uint32_t row_of(const uint32_t (&rows)[MAX_STAGES], int s) {
  switch (s) {
    case 0: return rows[0];
    case 1: return rows[1];
    case 2: return rows[2];
    case 3: return rows[3];
    case 4: return rows[4];
    case 5: return rows[5];
    case 6: return rows[6];
    default: return rows[7];
  }
}
__device__ __forceinline__ void set_row(uint32_t (&rows)[MAX_STAGES], int s, uint32_t v) {
  switch (s) {
    case 0: rows[0] = v; break;
    case 1: rows[1] = v; break;
    case 2: rows[2] = v; break;
    case 3: rows[3] = v; break;
    case 4: rows[4] = v; break;
    case 5: rows[5] = v; break;
    case 6: rows[6] = v; break;
    default: rows[7] = v; break;
  }
}

// Rust `f as i64` / `f as i32`: saturating, NaN -> 0.
__device__ __forceinline__ int64_t sat_i64(double x) {
  if (x != x) return 0;
  if (x >= 9223372036854775808.0) return INT64_MAX;
  if (x <= -9223372036854775808.0) return INT64_MIN;
  return (int64_t)x;
}
__device__ __forceinline__ int32_t sat_i32(double x) {
  if (x != x) return 0;
  if (x >= 2147483647.0) return INT32_MAX;
  if (x <= -2147483648.0) return INT32_MIN;
  return (int32_t)x;
}
__device__ __forceinline__ uint64_t sx32(uint32_t x) { return (uint64_t)(int64_t)(int32_t)x; }

// ----------------------------------------------------------------- interpreter
// Stack slots live in registers: every access below indexes r[] with a
// constant inside a switch on the (uniform) depth, so nothing spills.
__device__ __forceinline__ uint64_t get_slot(const uint64_t (&r)[MAX_DEPTH], int d) {
  switch (d) {
    case 0: return r[0];
    case 1: return r[1];
    case 2: return r[2];
    case 3: return r[3];
    case 4: return r[4];
    default: return r[5];
  }
}
__device__ __forceinline__ void set_slot(uint64_t (&r)[MAX_DEPTH], int d, uint64_t v) {
  switch (d) {
    case 0: r[0] = v; break;
    case 1: r[1] = v; break;
    case 2: r[2] = v; break;
    case 3: r[3] = v; break;
    case 4: r[4] = v; break;
    default: r[5] = v; break;
  }
}

// Values of up to two stage-0 columns already loaded for the current row
// (the wave-queued pipeline loads them for several rows at once); a LOAD of
// one of these columns takes the value from here instead of memory.
struct Pf {
  const void* c0 = nullptr;
  const void* c1 = nullptr;
  uint64_t v0 = 0, v1 = 0;
};

// Column value of a load kind (OP_LOAD_*) at a row.
__device__ __forceinline__ uint64_t load_col(uint32_t lop, uint64_t imm, uint64_t row) {
  if (lop == OP_LOAD_I32) return sx32((uint32_t)gp<int32_t>((const void*)imm)[row]);
  if (lop == OP_LOAD_U32) return (uint64_t)gp<uint32_t>((const void*)imm)[row];
  if (lop == OP_LOAD_I64) return gp<uint64_t>((const void*)imm)[row];
  return gp<uint8_t>((const void*)imm)[row] != 0;
}

// OP_CMPC: column value x compared with the constant i2.
__device__ __forceinline__ uint64_t cmpc(uint32_t flags, uint64_t x, uint64_t i2) {
  const uint32_t c = (flags >> 8) & 7u;
  bool r;
  if (flags & 0x800u) {
    const double a = u2f(x), b = u2f(i2);
    r = c == 0 ? a == b : c == 1 ? a != b : c == 2 ? a < b : c == 3 ? a <= b : c == 4 ? a > b : a >= b;
  } else {
    const int64_t a = (int64_t)x, b = (int64_t)i2;
    r = c == 0 ? a == b : c == 1 ? a != b : c == 2 ? a < b : c == 3 ? a <= b : c == 4 ? a > b : a >= b;
  }
  return r;
}

// Unary instruction on value a (validity in vres, may be cleared).
__device__ __forceinline__ uint64_t un_op(uint32_t op, uint64_t a, uint64_t imm, uint32_t& vres) {
  switch (op) {
    case OP_NOT: return a ^ 1ull;
    case OP_NEG_I32: return sx32(0u - (uint32_t)a);
    case OP_NEG_I64: return 0ull - a;
    case OP_NEG_F64: return a ^ 0x8000000000000000ull;
    case OP_I2F: return f2u((double)(int64_t)a);
    case OP_F2I64: return (uint64_t)sat_i64(u2f(a));
    case OP_F2I32: return sx32((uint32_t)sat_i32(u2f(a)));
    case OP_I64_TO_I32: return sx32((uint32_t)a);
    case OP_LOOKUP_U8: return gp<uint8_t>((const void*)imm)[a];
    case OP_LOOKUP_U64: return gp<uint64_t>((const void*)imm)[a];
    case OP_KEY_F64:
      if (a == 0x8000000000000000ull) return 0ull;
      if (u2f(a) != u2f(a))
        return imm ? NAN_KEY_TAG | (atomicAdd((unsigned long long*)imm, 1ull) & NAN_KEY_IDS) : CANONICAL_NAN;
      return a;
    case OP_JKEY_F64:
      if (u2f(a) != u2f(a)) vres = 0;
      return a == 0x8000000000000000ull ? 0ull : a;
    default: return u2f(a) == 0.0 ? 0ull : a;  // OP_HASH_F64
  }
}

// Binary instruction on (a, b) with validities va, vb; result validity in vres.
__device__ __forceinline__ uint64_t bin_op(uint32_t op, uint64_t a, uint64_t b, uint32_t va, uint32_t vb,
                                           uint32_t& vres) {
  vres = va & vb;
  switch (op) {
    case OP_ADD_I32: return sx32((uint32_t)a + (uint32_t)b);
    case OP_SUB_I32: return sx32((uint32_t)a - (uint32_t)b);
    case OP_MUL_I32: return sx32((uint32_t)a * (uint32_t)b);
    case OP_MOD_I32: {  // 0 and NULL for x % 0; wrapping_rem for -1
      const int32_t x = (int32_t)a, y = (int32_t)b;
      if (y == 0) vres = 0;
      return (y == 0 || y == -1) ? 0ull : sx32((uint32_t)(x % y));
    }
    case OP_ADD_I64: return a + b;
    case OP_SUB_I64: return a - b;
    case OP_MUL_I64: return a * b;
    case OP_MOD_I64: {
      const int64_t x = (int64_t)a, y = (int64_t)b;
      if (y == 0) vres = 0;
      return (y == 0 || y == -1) ? 0ull : (uint64_t)(x % y);
    }
    case OP_ADD_F64: return f2u(u2f(a) + u2f(b));
    case OP_SUB_F64: return f2u(u2f(a) - u2f(b));
    case OP_MUL_F64: return f2u(u2f(a) * u2f(b));
    case OP_DIV_F64:  // x / 0 is NULL (crates/minilake-exec/src/expr/eval.rs:260-270)
      if (u2f(b) == 0.0) vres = 0;
      return f2u(u2f(a) / u2f(b));
    case OP_MOD_F64: return f2u(fmod(u2f(a), u2f(b)));
    case OP_EQ_I: return (int64_t)a == (int64_t)b;
    case OP_NE_I: return (int64_t)a != (int64_t)b;
    case OP_LT_I: return (int64_t)a < (int64_t)b;
    case OP_LE_I: return (int64_t)a <= (int64_t)b;
    case OP_GT_I: return (int64_t)a > (int64_t)b;
    case OP_GE_I: return (int64_t)a >= (int64_t)b;
    case OP_EQ_F: return u2f(a) == u2f(b);
    case OP_NE_F: return u2f(a) != u2f(b);
    case OP_LT_F: return u2f(a) < u2f(b);
    case OP_LE_F: return u2f(a) <= u2f(b);
    case OP_GT_F: return u2f(a) > u2f(b);
    case OP_GE_F: return u2f(a) >= u2f(b);
    case OP_AND:  // Kleene AND (crates/minilake-exec/src/kernels/boolean.rs:18-36)
      vres = (va & vb) | (va & (uint32_t)(a == 0)) | (vb & (uint32_t)(b == 0));
      return a & b;
    default:  // OP_OR, Kleene (boolean.rs:40-58)
      vres = (va & vb) | (va & (uint32_t)(a != 0)) | (vb & (uint32_t)(b != 0));
      return a | b;
  }
}

// Evaluate a program for one row; returns validity, value in `val`.
__device__ __forceinline__ bool eval(const Code& ins, Prog p, const uint32_t (&rows)[MAX_STAGES], uint64_t& val,
                                     const Pf& pf = Pf{}) {
  const int first = uni(p.first), count = uni(p.count), nullable = uni(p.nullable);
  if (count == 1) {
    // a plain column load (most join keys, group keys and materialized
    // columns): no decode loop
    CInstr& in = ins.at(first);
    const uint64_t head = *(CU64*)&in;
    const uint32_t op = (uint32_t)head & 0xFFFFu;
    if (op < OP_CONST) {
      const int stage = (int)((head >> 24) & 0xFFu);
      const uint64_t imm = in.imm;
      const uint64_t row = row_of(rows, stage);
      if (stage == 0 && pf.c0 && (const void*)imm == pf.c0) val = pf.v0;
      else if (stage == 0 && pf.c1 && (const void*)imm == pf.c1) val = pf.v1;
      else val = load_col(op, imm, row);
      const uint64_t i2 = in.imm2;
      return !(nullable && i2) || gp<uint8_t>((const void*)i2)[row] != 0;
    }
  }
  uint64_t r[MAX_DEPTH];
#pragma unroll
  for (int i = 0; i < MAX_DEPTH; i++) r[i] = 0;
  uint32_t ok = 0xFFFFFFFFu;
#pragma unroll 1
  for (int pc = 0; pc < count; pc++) {
    CInstr& in = ins.at(first + pc);
    // op (u16) | depth (u8) << 16 | stage (u8) << 24 | flags << 32, one scalar load
    const uint64_t head = *(CU64*)&in;
    const uint32_t op = (uint32_t)head & 0xFFFFu;
    const int d = (int)((head >> 16) & 0xFFu);
    const int stage = (int)((head >> 24) & 0xFFu);
    const uint32_t flags = (uint32_t)(head >> 32);
    const uint64_t imm = in.imm;
    uint64_t res;
    uint32_t vres = 1;
    int dst;
    if (op <= OP_LOOKUPC) {  // push
      dst = d;
      if (op == OP_CONST) {
        res = imm;
        vres = flags & 1u;
      } else {
        const uint64_t row = row_of(rows, stage);
        const uint32_t lop = column_load_op(op, flags);
        uint64_t x;
        if (stage == 0 && pf.c0 && (const void*)imm == pf.c0) x = pf.v0;
        else if (stage == 0 && pf.c1 && (const void*)imm == pf.c1) x = pf.v1;
        else x = load_col(lop, imm, row);
        const uint64_t i2 = in.imm2;
        if (op == OP_CMPC) {
          res = cmpc(flags, x, i2);
        } else if (op == OP_LOOKUPC) {
          res = gp<uint8_t>((const void*)i2)[x];
        } else {
          res = x;
          if (nullable && i2) vres = gp<uint8_t>((const void*)i2)[row] != 0;
        }
      }
    } else if (op < OP_ADD_I32) {  // unary on the top slot
      dst = d - 1;
      vres = (ok >> dst) & 1u;
      res = un_op(op, get_slot(r, dst), imm, vres);
    } else {  // binary: slots d-2 (left), d-1 (right) -> d-2
      dst = d - 2;
      res = bin_op(op, get_slot(r, dst), get_slot(r, d - 1), (ok >> dst) & 1u, (ok >> (d - 1)) & 1u, vres);
    }
    set_slot(r, dst, res);
    ok = (ok & ~(1u << dst)) | (vres << dst);
  }
  val = r[0];
  return (ok & 1u) != 0;
}

// The same program over 4 stage-0 rows per lane (src[u]). Each instruction
// is decoded once for the 4 rows: one switch on its stack slot and one on its
// opcode, with the rows handled inside the case, so the scalar unit (which
// runs the interpreter's decode and branches for every wave of a CU) does the
// same work for 4 rows as for one. The stack is 4 named vector registers
// (no array, so nothing lands in scratch memory); the host runs a sink this
// way only when every program it evaluates needs at most 4 slots
// (Plan::src_batch). Column loads are issued for all 4 rows before use.
typedef uint64_t u64x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

struct Stack4 {
  u64x4 s0, s1, s2, s3;
  u32x4 ok;  // bit k: slot k valid
};

// Slot selection by switch (whose cases copy named registers; branch trees
// of ternaries are folded into selects of addresses, which puts the stack in
// scratch memory); the opcode dispatch tests the hot cases first, since a
// switch on a uniform value is lowered to a chain of compare-and-branch tests.
__device__ __forceinline__ u64x4 st_get(const Stack4& S, int k) {
  switch (k) {
    case 0: return S.s0;
    case 1: return S.s1;
    case 2: return S.s2;
    default: return S.s3;
  }
}
__device__ __forceinline__ void st_get2(const Stack4& S, int k, u64x4& a, u64x4& b) {
  switch (k) {
    case 0: a = S.s0; b = S.s1; break;
    case 1: a = S.s1; b = S.s2; break;
    default: a = S.s2; b = S.s3; break;
  }
}
__device__ __forceinline__ void st_set(Stack4& S, int k, const u64x4& v) {
  switch (k) {
    case 0: S.s0 = v; break;
    case 1: S.s1 = v; break;
    case 2: S.s2 = v; break;
    default: S.s3 = v; break;
  }
}

// The staged columns of a round (Plan::sp_col) for the lane's 4 rows.
struct Staged8 {
  u64x4 c0, c1, c2, c3, c4, c5, c6, c7;
};
__device__ __forceinline__ u64x4 load4(uint32_t lop, const void* col, const uint32_t (&src)[4]) {
  if (lop == OP_LOAD_I64) {
    const auto c = gp<uint64_t>(col);
    return (u64x4){c[src[0]], c[src[1]], c[src[2]], c[src[3]]};
  }
  if (lop == OP_LOAD_I32) {
    const auto c = gp<int32_t>(col);
    return (u64x4){sx32((uint32_t)c[src[0]]), sx32((uint32_t)c[src[1]]), sx32((uint32_t)c[src[2]]),
                   sx32((uint32_t)c[src[3]])};
  }
  if (lop == OP_LOAD_U32) {
    const auto c = gp<uint32_t>(col);
    return (u64x4){(uint64_t)c[src[0]], (uint64_t)c[src[1]], (uint64_t)c[src[2]], (uint64_t)c[src[3]]};
  }
  const auto c = gp<uint8_t>(col);
  return (u64x4){(uint64_t)(c[src[0]] != 0), (uint64_t)(c[src[1]] != 0), (uint64_t)(c[src[2]] != 0),
                 (uint64_t)(c[src[3]] != 0)};
}
// Every staged column's loads are issued before any value is used.
__device__ __forceinline__ void stage_load(const Plan& P, const uint32_t (&src)[4], Staged8& T) {
  const int n = P.sp_n;
  if (n > 0) T.c0 = load4(P.sp_op[0], P.sp_col[0], src);
  if (n > 1) T.c1 = load4(P.sp_op[1], P.sp_col[1], src);
  if (n > 2) T.c2 = load4(P.sp_op[2], P.sp_col[2], src);
  if (n > 3) T.c3 = load4(P.sp_op[3], P.sp_col[3], src);
  if (n > 4) T.c4 = load4(P.sp_op[4], P.sp_col[4], src);
  if (n > 5) T.c5 = load4(P.sp_op[5], P.sp_col[5], src);
  if (n > 6) T.c6 = load4(P.sp_op[6], P.sp_col[6], src);
  if (n > 7) T.c7 = load4(P.sp_op[7], P.sp_col[7], src);
}
__device__ __forceinline__ u64x4 staged(const Staged8& T, uint32_t j) {
  switch (j) {
    case 0: return T.c0;
    case 1: return T.c1;
    case 2: return T.c2;
    case 3: return T.c3;
    case 4: return T.c4;
    case 5: return T.c5;
    case 6: return T.c6;
    default: return T.c7;
  }
}

// NUL: the program can produce NULL (Prog::nullable); without it every value
// is valid, so no validity is loaded or tracked.
// start > 0: the first `start` instructions are those of a program already
// evaluated, whose result (val, valid) becomes stack slot 0.
template <bool NUL>
__device__ __forceinline__ void eval_src4_t(const Code& ins, Prog p, int start, const uint32_t (&src)[4],
                                            const Staged8& T, uint64_t (&val)[4], bool (&valid)[4]) {
  Stack4 S;
  S.s0 = S.s1 = S.s2 = S.s3 = (u64x4)(0);
  S.ok = (u32x4)(0xFFFFFFFFu);
  if (start > 0) {
    S.s0 = (u64x4){val[0], val[1], val[2], val[3]};
    if (NUL)
      S.ok = (u32x4){valid[0] ? 0xFFFFFFFFu : 0xFFFFFFFEu, valid[1] ? 0xFFFFFFFFu : 0xFFFFFFFEu,
                     valid[2] ? 0xFFFFFFFFu : 0xFFFFFFFEu, valid[3] ? 0xFFFFFFFFu : 0xFFFFFFFEu};
  }
  const int first = uni(p.first), count = uni(p.count);
#pragma unroll 1
  for (int pc = start; pc < count; pc++) {
    CInstr& in = ins.at(first + pc);
    const uint64_t head = *(CU64*)&in;
    const uint32_t op = (uint32_t)head & 0xFFFFu;
    const int d = (int)((head >> 16) & 0xFFu);
    const uint32_t flags = (uint32_t)(head >> 32);
    const uint64_t imm = in.imm;
    u64x4 res;
    u32x4 vres = (u32x4)(1u);
    int dst;
    if (op <= OP_LOOKUPC) {
      dst = d;
      if (op == OP_CONST) {
        res = (u64x4)(imm);
        vres = (u32x4)(flags & 1u);
      } else {
        const uint32_t lop = column_load_op(op, flags);
        const uint64_t i2 = in.imm2;
        const uint32_t sj = (flags >> 24) & 15u;
        const u64x4 x = sj ? staged(T, sj - 1u) : load4(lop, (const void*)imm, src);
        if (op == OP_CMPC) {
          res = (u64x4){cmpc(flags, x[0], i2), cmpc(flags, x[1], i2), cmpc(flags, x[2], i2), cmpc(flags, x[3], i2)};
        } else if (op == OP_LOOKUPC) {
          const auto t = gp<uint8_t>((const void*)i2);
          res = (u64x4){(uint64_t)t[x[0]], (uint64_t)t[x[1]], (uint64_t)t[x[2]], (uint64_t)t[x[3]]};
        } else {
          res = x;
          if (NUL && i2) {
            const auto vb = gp<uint8_t>((const void*)i2);
            vres = (u32x4){(uint32_t)(vb[src[0]] != 0), (uint32_t)(vb[src[1]] != 0), (uint32_t)(vb[src[2]] != 0),
                           (uint32_t)(vb[src[3]] != 0)};
          }
        }
      }
    } else if (op < OP_ADD_I32) {
      dst = d - 1;
      const u64x4 a = st_get(S, dst);
      if (NUL) vres = (S.ok >> (uint32_t)dst) & 1u;
      uint32_t v0 = vres[0], v1 = vres[1], v2 = vres[2], v3 = vres[3];
#define ML_UN(OPC) res = (u64x4){un_op(OPC, a[0], imm, v0), un_op(OPC, a[1], imm, v1), un_op(OPC, a[2], imm, v2), un_op(OPC, a[3], imm, v3)};
      if (op == OP_I2F) {
        ML_UN(OP_I2F)
      } else if (op == OP_NEG_F64) {
        ML_UN(OP_NEG_F64)
      } else {
        switch (op) {
#define ML_UNC(OPC) case OPC: ML_UN(OPC) break;
          ML_UNC(OP_NOT) ML_UNC(OP_NEG_I32) ML_UNC(OP_NEG_I64) ML_UNC(OP_F2I64) ML_UNC(OP_F2I32)
          ML_UNC(OP_I64_TO_I32) ML_UNC(OP_LOOKUP_U8) ML_UNC(OP_LOOKUP_U64) ML_UNC(OP_KEY_F64) ML_UNC(OP_JKEY_F64)
          default: ML_UN(OP_HASH_F64) break;
#undef ML_UNC
        }
      }
#undef ML_UN
      vres = (u32x4){v0, v1, v2, v3};
    } else {
      dst = d - 2;
      u64x4 a, b;
      st_get2(S, dst, a, b);
      const u32x4 va = NUL ? (S.ok >> (uint32_t)dst) & 1u : (u32x4)(1u), vb = NUL ? (S.ok >> (uint32_t)(dst + 1)) & 1u : (u32x4)(1u);
      uint32_t v0 = 1, v1 = 1, v2 = 1, v3 = 1;
#define ML_BIN(OPC) res = (u64x4){bin_op(OPC, a[0], b[0], va[0], vb[0], v0), bin_op(OPC, a[1], b[1], va[1], vb[1], v1), bin_op(OPC, a[2], b[2], va[2], vb[2], v2), bin_op(OPC, a[3], b[3], va[3], vb[3], v3)};
#define ML_BINC(OPC) case OPC: ML_BIN(OPC) break;
      if (op == OP_MUL_F64) {
        ML_BIN(OP_MUL_F64)
      } else if (op == OP_SUB_F64) {
        ML_BIN(OP_SUB_F64)
      } else if (op == OP_ADD_F64) {
        ML_BIN(OP_ADD_F64)
      } else if (op < OP_EQ_I) {  // the other arithmetic
        switch (op) {
          ML_BINC(OP_ADD_I64) ML_BINC(OP_SUB_I64) ML_BINC(OP_MUL_I64) ML_BINC(OP_DIV_F64) ML_BINC(OP_ADD_I32)
          ML_BINC(OP_SUB_I32) ML_BINC(OP_MUL_I32) ML_BINC(OP_MOD_I32) ML_BINC(OP_MOD_I64)
          default: ML_BIN(OP_MOD_F64) break;
        }
      } else if (op < OP_EQ_F) {  // integer comparisons
        switch (op) {
          ML_BINC(OP_EQ_I) ML_BINC(OP_NE_I) ML_BINC(OP_LT_I) ML_BINC(OP_LE_I) ML_BINC(OP_GT_I)
          default: ML_BIN(OP_GE_I) break;
        }
      } else {  // float comparisons, AND, OR
        switch (op) {
          ML_BINC(OP_EQ_F) ML_BINC(OP_NE_F) ML_BINC(OP_LT_F) ML_BINC(OP_LE_F) ML_BINC(OP_GT_F) ML_BINC(OP_GE_F)
          ML_BINC(OP_AND)
          default: ML_BIN(OP_OR) break;
        }
      }
#undef ML_BINC
#undef ML_BIN
      vres = (u32x4){v0, v1, v2, v3};
    }
    st_set(S, dst, res);
    if (NUL) S.ok = (S.ok & ~(1u << (uint32_t)dst)) | (vres << (uint32_t)dst);
  }
#pragma unroll
  for (int u = 0; u < 4; u++) {
    val[u] = S.s0[u];
    valid[u] = !NUL || (S.ok[u] & 1u) != 0;
  }
}

__device__ __forceinline__ void eval_src4(const Code& ins, Prog p, const uint32_t (&src)[4], const Staged8& T,
                                          uint64_t (&val)[4], bool (&valid)[4], int start = 0) {
  if (uni(p.nullable)) eval_src4_t<true>(ins, p, start, src, T, val, valid);
  else eval_src4_t<false>(ins, p, start, src, T, val, valid);
}

__device__ __forceinline__ void set_word(uint64_t (&w)[MAX_KEYS + 1], int k, uint64_t v) {
  switch (k) {
    case 0: w[0] = v; break;
    case 1: w[1] = v; break;
    case 2: w[2] = v; break;
    case 3: w[3] = v; break;
    default: w[4] = v; break;
  }
}

// Table hash of key words (n is uniform; the unrolled body keeps w in registers).
__device__ __forceinline__ uint64_t key_hash(const uint64_t* w, int n) {
  uint64_t h = tmix(w[0] ^ TSEED);
#pragma unroll
  for (int k = 1; k < MAX_KEYS + 1; k++)
    if (k < n) h = tmix(h ^ w[k]);
  return h;
}

__device__ __forceinline__ bool bloom_maybe(const DevJoinTable& T, uint64_t h) {
  if (!T.bloom) return true;
  const uint64_t g = bloom_mix(h);
  const uint64_t bits = bloom_bits(g);
  return (T.bloom[bloom_word(T, h, g)] & bits) == bits;
}

// Dense table lookup of key word k: presence bit, then the row: the rank of
// the key's bit (a rank directory), or its offset itself (rows held at their
// key offset, no directory).
__device__ __forceinline__ bool dense_find(const DevJoinTable& T, uint64_t k, uint32_t& row) {
  const uint64_t idx = k - (uint64_t)T.dense_min;
  if (idx >= T.dense_range) return false;
  if (T.dense_full) {
    row = (uint32_t)idx;
    return true;
  }
  const uint64_t bits = gp<uint64_t>(T.dense_bits)[idx >> 6];
  const uint32_t b = (uint32_t)(idx & 63);
  if (!((bits >> b) & 1ull)) return false;
  row = T.dense_rank ? gp<uint32_t>(T.dense_rank)[idx >> 6] + (uint32_t)__popcll(bits & ((1ull << b) - 1ull))
                     : (uint32_t)idx;
  return true;
}

// First slot at or after `pos` whose key equals w; false at the first empty slot.
__device__ __forceinline__ bool join_next(const DevJoinTable& T, const uint64_t (&w)[MAX_KEYS + 1], uint64_t h,
                                          uint64_t& pos, int64_t& row) {
  const int stride = 1 + T.nkw;
  const uint32_t tag = (uint32_t)(h >> 32);
  for (;;) {
    const uint64_t* s = T.slots + pos * stride;
    const uint64_t hdr = s[0];
    if (hdr == 0) return false;
    if ((uint32_t)(hdr >> 32) == tag) {
      bool eq = true;
#pragma unroll
      for (int k = 0; k < MAX_KEYS; k++)
        if (k < T.nkw) eq = eq && (s[1 + k] == w[k]);
      if (eq) {
        row = (int64_t)(uint32_t)hdr - 1;
        return true;
      }
    }
    pos = join_step(T, pos);
  }
}

__device__ __forceinline__ bool probe_keys(const Plan& P, const Code& ins, const Step& st,
                                           const uint32_t (&rows)[MAX_STAGES], uint64_t (&w)[MAX_KEYS + 1],
                                           const Pf& pf = Pf{}) {
  bool ok = true;
#pragma unroll
  for (int k = 0; k < MAX_KEYS + 1; k++) w[k] = 0;
  if (uni(st.nkeys) == 2 && uni(st.keys[0].count) == 1 && uni(st.keys[1].count) == 1 && !pf.c0) {
    // two keys, each a plain column load (any width, as the single-load path
    // of eval): both loads issued before either is used
    CInstr& i0 = ins.at(st.keys[0].first);
    CInstr& i1 = ins.at(st.keys[1].first);
    const uint64_t h0 = *(CU64*)&i0, h1 = *(CU64*)&i1;
    const uint32_t op0 = (uint32_t)h0 & 0xFFFFu, op1 = (uint32_t)h1 & 0xFFFFu;
    if (op0 < OP_CONST && op1 < OP_CONST) {
      const uint64_t r0 = row_of(rows, (int)((h0 >> 24) & 0xFFu)), r1 = row_of(rows, (int)((h1 >> 24) & 0xFFu));
      const uint64_t a = load_col(op0, i0.imm, r0), b = load_col(op1, i1.imm, r1);
      const uint64_t v0 = uni(st.keys[0].nullable) ? i0.imm2 : 0, v1 = uni(st.keys[1].nullable) ? i1.imm2 : 0;
      const bool ok0 = !v0 || gp<uint8_t>((const void*)v0)[r0] != 0;
      const bool ok1 = !v1 || gp<uint8_t>((const void*)v1)[r1] != 0;
      set_word(w, 0, a);
      set_word(w, 1, b);
      return ok0 && ok1;
    }
  }
#pragma unroll 1
  for (int k = 0; k < st.nkeys; k++) {
    uint64_t v;
    ok = eval(ins, st.keys[k], rows, v, pf) && ok;
    set_word(w, k, v);
  }
  return ok;
}

// Filters and unique probes in [from, to); false once the row is dropped.
__device__ __forceinline__ bool run_steps(const Plan& P, const Code& ins, int from, int to,
                                          uint32_t (&rows)[MAX_STAGES], const Pf& pf = Pf{}) {
  bool alive = true;
  for (int s_ = from; s_ < to; s_++) {
    if (!alive) break;
    const int s = min(uni(s_), MAX_STEPS - 1);
    const Step& st = P.steps[s];
    if (st.kind == STEP_FILTER) {
      uint64_t v;
      const bool ok = eval(ins, st.pred, rows, v, pf);
      alive = ok && v != 0;
    } else {
      uint64_t w[MAX_KEYS + 1];
      if (!probe_keys(P, ins, st, rows, w, pf)) {  // NULL keys never match (join.rs:284-286)
        alive = false;
        continue;
      }
      const DevJoinTable& T = P.tables[st.table];
      if (T.dense) {
        uint32_t drow;
        bool ok = true;
        if (T.bloom) {
          const uint64_t i = w[0] - (uint64_t)T.dense_min;
          const uint64_t bits = bloom_bits(bloom_mix(key_hash(w, T.nkw)));
          ok = i < T.dense_range && (gp<uint64_t>(T.bloom)[i >> DENSE_BLOOM_SHIFT] & bits) == bits;
        }
        ok = ok && dense_find(T, w[0], drow);
        for (int k = 1; k < T.nkw && ok; k++) ok = dense_key_word(T, drow, k) == w[k];
        if (!ok) {
          alive = false;
          continue;
        }
        set_row(rows, st.stage, drow);
        continue;
      }
      const uint64_t h = key_hash(w, st.nkeys);
      uint64_t pos = h & T.mask;
      int64_t row = 0;
      if (!bloom_maybe(T, h) || !join_next(T, w, h, pos, row)) {
        alive = false;
        continue;
      }
      set_row(rows, st.stage, row);
    }
  }
  return alive;
}

// ------------------------------------------------------------- wave helpers
__device__ __forceinline__ uint64_t wave_alloc(uint64_t* counter, uint32_t n_per_lane) {
  // one atomic per wave for the active lanes; each lane gets its own range
  const uint64_t active = __ballot(1);
  const int lane = __lane_id();
  const int leader = __ffsll((unsigned long long)active) - 1;
  const uint64_t below = active & ((1ull << lane) - 1ull);
  const uint32_t rank = (uint32_t)__popcll(below);
  const uint32_t total = (uint32_t)__popcll(active) * n_per_lane;
  uint64_t base = 0;
  if (lane == leader) base = atomicAdd((unsigned long long*)counter, (unsigned long long)total);
  base = __shfl(base, leader);
  return base + (uint64_t)rank * n_per_lane;
}

// --------------------------------------------------------------- aggregates
__device__ __forceinline__ void atomic_min_f64(double* p, double v) {
  unsigned long long* q = (unsigned long long*)p;
  unsigned long long cur = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  while (v < __longlong_as_double((long long)cur)) {
    unsigned long long prev = atomicCAS(q, cur, (unsigned long long)__double_as_longlong(v));
    if (prev == cur) break;
    cur = prev;
  }
}
__device__ __forceinline__ void atomic_max_f64(double* p, double v) {
  unsigned long long* q = (unsigned long long*)p;
  unsigned long long cur = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  while (v > __longlong_as_double((long long)cur)) {
    unsigned long long prev = atomicCAS(q, cur, (unsigned long long)__double_as_longlong(v));
    if (prev == cur) break;
    cur = prev;
  }
}

// fold state b into state a
__device__ __forceinline__ uint64_t acc_combine(int kind, uint64_t a, uint64_t b) {
  switch (kind) {
    case ACC_SUM_F64: return f2u(u2f(a) + u2f(b));
    case ACC_SUM_I64: case ACC_COUNT: return a + b;
    case ACC_MIN_I64: return (int64_t)b < (int64_t)a ? b : a;
    case ACC_MAX_I64: return (int64_t)b > (int64_t)a ? b : a;
    case ACC_MIN_F64: return u2f(b) < u2f(a) ? b : a;
    case ACC_MAX_F64: return u2f(b) > u2f(a) ? b : a;
    default: return a;
  }
}

// accumulate into a global (hash table) state slot
__device__ __forceinline__ void acc_apply_global(int kind, void* data, int64_t slot, uint64_t v) {
  switch (kind) {
    case ACC_SUM_F64: unsafeAtomicAdd((double*)gpw<double>(data) + slot, u2f(v)); break;
    case ACC_SUM_I64: atomicAdd((unsigned long long*)gpw<uint64_t>(data) + slot, (unsigned long long)v); break;
    case ACC_COUNT: atomicAdd((unsigned long long*)gpw<uint64_t>(data) + slot, 1ull); break;
    case ACC_MIN_I64: atomicMin((long long*)gpw<uint64_t>(data) + slot, (long long)v); break;
    case ACC_MAX_I64: atomicMax((long long*)gpw<uint64_t>(data) + slot, (long long)v); break;
    case ACC_MIN_F64: atomic_min_f64((double*)data + slot, u2f(v)); break;
    case ACC_MAX_F64: atomic_max_f64((double*)data + slot, u2f(v)); break;
    default: break;
  }
}

// accumulate into an LDS state word
__device__ __forceinline__ void acc_apply_lds(int kind, uint64_t* w, uint64_t v) {
  switch (kind) {
    case ACC_SUM_F64: unsafeAtomicAdd((double*)w, u2f(v)); break;
    case ACC_SUM_I64: atomicAdd((unsigned long long*)w, (unsigned long long)v); break;
    case ACC_COUNT: atomicAdd((unsigned long long*)w, 1ull); break;
    case ACC_MIN_I64: atomicMin((long long*)w, (long long)v); break;
    case ACC_MAX_I64: atomicMax((long long*)w, (long long)v); break;
    case ACC_MIN_F64: {
      unsigned long long* q = (unsigned long long*)w;
      unsigned long long cur = *q;
      while (u2f(v) < __longlong_as_double((long long)cur)) {
        unsigned long long prev = atomicCAS(q, cur, (unsigned long long)v);
        if (prev == cur) break;
        cur = prev;
      }
      break;
    }
    case ACC_MAX_F64: {
      unsigned long long* q = (unsigned long long*)w;
      unsigned long long cur = *q;
      while (u2f(v) > __longlong_as_double((long long)cur)) {
        unsigned long long prev = atomicCAS(q, cur, (unsigned long long)v);
        if (prev == cur) break;
        cur = prev;
      }
      break;
    }
    default: break;
  }
}

// Wait until this thread's earlier stores have completed. After agent-scope
// atomic stores (sc1) this orders them before the thread's later stores at
// agent scope, which is what an agent-scope release does for atomic stores;
// the release would also write back the whole L2 (buffer_wbl2), which it
// needs only for non-atomic stores and which costs ~20x more per publication.
__device__ __forceinline__ void wait_for_stores() {
  __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0); expcnt and lgkmcnt not waited on
}

// Find or insert the group with key words w; returns its slot or -1 (table full).
// Slot words are read with relaxed agent-scope loads, which bypass the CU's
// non-coherent L1: an acquire load would add a full wait and an L1 invalidate
// to every probe. The inserter stores the keys as agent-scope atomics, waits
// for them, then stores the header; a reader compares keys only after its
// header load returned that header (the key loads are issued under that
// branch), so the keys it reads are the inserted ones.
// Returns -1 when the row found no slot, -2 when an earlier row of a
// pre-sized table already found none (the batch is rerun: nothing to record).
__device__ __forceinline__ int64_t agg_slot(const Plan& P, const uint64_t (&w)[MAX_KEYS + 1], int nkw, uint64_t h) {
  const uint64_t hdr = (h & ~3ull) | 2ull;
  const uint64_t stride = 1 + (uint64_t)nkw;
  uint64_t pos = h & P.agg_mask;
  uint64_t probes = 0;
  if (P.agg_presized && (__hip_atomic_load(P.agg_overflow, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & 2u))
    return -2;
  while (probes <= P.agg_wrap) {
    uint64_t* s = P.agg_slots + pos * stride;
    uint64_t cur = __hip_atomic_load(s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (cur == 0) {
      const uint64_t prev = atomicCAS((unsigned long long*)s, 0ull, 1ull);
      if (prev == 0) {
        const uint64_t n = atomicAdd((unsigned long long*)P.agg_count, 1ull);
        if (n >= P.agg_fill_limit) atomicOr(P.agg_overflow, P.agg_presized ? 3u : 1u);  // past the load-factor limit
#pragma unroll
        for (int k = 0; k < MAX_KEYS + 1; k++)
          if (k < nkw) __hip_atomic_store(s + 1 + k, w[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        wait_for_stores();
        __hip_atomic_store(s, hdr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return (int64_t)pos;
      }
      cur = prev;
      if (cur != 1) cur = __hip_atomic_load(s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (cur == 1) continue;  // another lane is writing this slot's key: read it again
    if (cur == hdr) {
      bool eq = true;
#pragma unroll
      for (int k = 0; k < MAX_KEYS + 1; k++)
        if (k < nkw) eq = eq && (__hip_atomic_load(s + 1 + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == w[k]);
      if (eq) return (int64_t)pos;
    }
    pos = (pos & ~P.agg_wrap) | ((pos + 1) & P.agg_wrap);
    probes++;
  }
  return -1;
}

__device__ __forceinline__ int agg_keys(const Plan& P, const Code& ins, const uint32_t (&rows)[MAX_STAGES],
                                        uint64_t (&w)[MAX_KEYS + 1]) {
  uint64_t nullmask = 0;
#pragma unroll
  for (int k = 0; k < MAX_KEYS + 1; k++) w[k] = 0;
#pragma unroll 1
  for (int k = 0; k < P.nkeys; k++) {
    uint64_t v;
    if (!eval(ins, P.keys[k], rows, v)) {  // NULL keys form one group (GROUP BY semantics)
      v = 0;
      nullmask |= 1ull << k;
    }
    set_word(w, k, v);
  }
  if (P.key_nullable) set_word(w, P.nkeys, nullmask);
  return P.nkeys + (P.key_nullable ? 1 : 0);
}

__device__ __forceinline__ void sink_hash_agg(const Plan& P, const Code& ins, const uint32_t (&rows)[MAX_STAGES]) {
  uint64_t w[MAX_KEYS + 1];
  const int nkw = agg_keys(P, ins, rows, w);
  const uint64_t h = key_hash(w, nkw);
  const int64_t slot = agg_slot(P, w, nkw, h);
  if (slot < 0) {  // table full: this row's contribution is lost, the caller must rerun
    if (slot == -1) atomicOr(P.agg_overflow, 2u);
    return;
  }
  for (int a = 0; a < P.nargs; a++) {
    uint64_t v;
    const bool ok = eval(ins, P.args[a], rows, v);
    if (!ok) continue;  // NULL arguments are skipped by every accumulator
    for (int s = 0; s < P.nacc; s++)
      if (P.acc[s].arg == a) acc_apply_global(P.acc[s].kind, P.acc_data[s], slot, v);
  }
  for (int s = 0; s < P.nacc; s++)
    if (P.acc[s].arg < 0) atomicAdd((unsigned long long*)P.acc_data[s] + slot, 1ull);
}

// Merge an accumulator state (not a row value) into a global state slot.
__device__ __forceinline__ void acc_merge_global(int kind, void* data, int64_t slot, uint64_t v) {
  switch (kind) {
    case ACC_SUM_F64: unsafeAtomicAdd((double*)gpw<double>(data) + slot, u2f(v)); break;
    case ACC_SUM_I64: case ACC_COUNT: atomicAdd((unsigned long long*)gpw<uint64_t>(data) + slot, (unsigned long long)v); break;
    case ACC_MIN_I64: atomicMin((long long*)gpw<uint64_t>(data) + slot, (long long)v); break;
    case ACC_MAX_I64: atomicMax((long long*)gpw<uint64_t>(data) + slot, (long long)v); break;
    case ACC_MIN_F64: atomic_min_f64((double*)gpw<double>(data) + slot, u2f(v)); break;
    case ACC_MAX_F64: atomic_max_f64((double*)gpw<double>(data) + slot, u2f(v)); break;
    default: break;
  }
}

// Hash aggregate through the workgroup's accumulator cache (ac: P.nacc arrays
// of P.ac_n states, then P.ac_n tags = slot + 1, 0 = free).
__device__ __forceinline__ void sink_hash_agg_cached(const Plan& P, const Code& ins, const uint32_t (&rows)[MAX_STAGES],
                                                     uint64_t* ac) {
  uint64_t w[MAX_KEYS + 1];
  const int nkw = agg_keys(P, ins, rows, w);
  const uint64_t h = key_hash(w, nkw);
  const int64_t slot = agg_slot(P, w, nkw, h);
  if (slot < 0) {
    if (slot == -1) atomicOr(P.agg_overflow, 2u);
    return;
  }
  uint32_t* tags = (uint32_t*)(ac + (uint64_t)P.ac_n * P.nacc);
  // a group's cache entry: the first of AC_PROBES entries from slot mod ac_n
  // that holds it or is free (claimed with a CAS); otherwise the global table
  constexpr int AC_PROBES = 4;
  const uint32_t want = (uint32_t)slot + 1u;
  uint32_t ci = (uint32_t)slot & (uint32_t)(P.ac_n - 1);
  bool cached = false;
  for (int pr = 0; pr < AC_PROBES; pr++) {
    uint32_t tag = tags[ci];
    if (tag == 0) {
      const uint32_t prev = atomicCAS(&tags[ci], 0u, want);
      tag = prev == 0 ? want : prev;
    }
    if (tag == want) {
      cached = true;
      break;
    }
    ci = (ci + 1u) & (uint32_t)(P.ac_n - 1);
  }
  for (int a = 0; a < P.nargs; a++) {
    uint64_t v;
    const bool ok = eval(ins, P.args[a], rows, v);
    if (!ok) continue;
    for (int s = 0; s < P.nacc; s++)
      if (P.acc[s].arg == a) {
        if (cached) acc_apply_lds(P.acc[s].kind, ac + (uint64_t)s * P.ac_n + ci, v);
        else acc_apply_global(P.acc[s].kind, P.acc_data[s], slot, v);
      }
  }
  for (int s = 0; s < P.nacc; s++)
    if (P.acc[s].arg < 0) {
      if (cached) atomicAdd((unsigned long long*)(ac + (uint64_t)s * P.ac_n + ci), 1ull);
      else atomicAdd((unsigned long long*)gpw<uint64_t>(P.acc_data[s]) + slot, 1ull);
    }
}

__device__ __forceinline__ void sink_direct_agg(const Plan& P, const Code& ins, const uint32_t (&rows)[MAX_STAGES],
                                                uint64_t* lds) {
  int64_t g = 0;
#pragma unroll 1
  for (int k = 0; k < P.nkeys; k++) {
    uint64_t v;
    const bool ok = eval(ins, P.keys[k], rows, v);
    // a NULL key takes the extra last value of its dimension
    const int64_t code = ok ? (int64_t)v : P.dim_size[k] - 1;
    g += code * P.dim_stride[k];
  }
  uint64_t* st = lds + g * P.nacc;
  for (int a = 0; a < P.nargs; a++) {
    uint64_t v;
    const bool ok = eval(ins, P.args[a], rows, v);
    if (!ok) continue;
    for (int s = 0; s < P.nacc; s++)
      if (P.acc[s].arg == a) acc_apply_lds(P.acc[s].kind, st + s, v);
  }
  for (int s = 0; s < P.nacc; s++)
    if (P.acc[s].arg < 0) atomicAdd((unsigned long long*)(st + s), 1ull);
}

// The direct aggregate for U source rows per lane (a level-0 round that ends
// the pipeline): keys and arguments evaluated once per instruction for all U
// rows, then each live row's states added to the lane's copy of the cells.
template <int U>
__device__ __forceinline__ void sink_direct_agg_src(const Plan& P, const Code& ins, const uint32_t (&src)[U],
                                                    const Staged8& T, const bool (&alive)[U], uint64_t* cells) {
  static_assert(U == 4, "the multi-row evaluator takes 4 rows per lane");
  int64_t g[U];
  uint64_t v[U];
  bool ok[U];
#pragma unroll
  for (int u = 0; u < U; u++) g[u] = 0;
#pragma unroll 1
  for (int k = 0; k < P.nkeys; k++) {
    eval_src4(ins, P.keys[k], src, T, v, ok);
    const int64_t stride = P.dim_stride[k], nullcode = P.dim_size[k] - 1;
#pragma unroll
    for (int u = 0; u < U; u++) g[u] += (ok[u] ? (int64_t)v[u] : nullcode) * stride;
  }
  const int nacc = P.nacc;
#pragma unroll
  for (int u = 0; u < U; u++) g[u] *= nacc;
  // the slots an argument feeds, from the host's per-argument masks
  auto apply = [&](uint32_t m, bool count) {
    while (m) {
      const int s = __builtin_ctz(m);
      m &= m - 1u;
      const int kind = count ? (int)ACC_COUNT : uni(P.acc[s].kind);
      bool on[U];
#pragma unroll
      for (int u = 0; u < U; u++) on[u] = alive[u] && (count || ok[u]);
#define ML_ACC(K) _Pragma("unroll") for (int u = 0; u < U; u++) if (on[u]) acc_apply_lds(K, cells + g[u] + s, v[u]);
      if (kind == ACC_SUM_F64) {
        ML_ACC(ACC_SUM_F64)
      } else if (kind == ACC_COUNT) {
        ML_ACC(ACC_COUNT)
      } else if (kind == ACC_SUM_I64) {
        ML_ACC(ACC_SUM_I64)
      } else if (kind == ACC_MIN_I64) {
        ML_ACC(ACC_MIN_I64)
      } else if (kind == ACC_MAX_I64) {
        ML_ACC(ACC_MAX_I64)
      } else if (kind == ACC_MIN_F64) {
        ML_ACC(ACC_MIN_F64)
      } else {
        ML_ACC(ACC_MAX_F64)
      }
#undef ML_ACC
    }
  };
#pragma unroll 1
  for (int a = 0; a < P.nargs; a++) {
    // an argument that begins with the previous one's code continues from
    // its values, still in v / ok
    const int from = ((uni(P.arg_chain) >> a) & 1u) ? uni(P.args[a - 1].count) : 0;
    eval_src4(ins, P.args[a], src, T, v, ok, from);
    apply(uni(P.arg_slots[a]), false);
  }
  apply(uni(P.count_slots), true);
}

// Materialize every column of the layout for one row at output position pos.
__device__ __forceinline__ void write_cols(const Plan& P, const Code& ins, const uint32_t (&rows)[MAX_STAGES], uint64_t pos) {
  for (int c = 0; c < P.ncols; c++) {
    uint64_t v;
    const bool ok = eval(ins, P.cols[c], rows, v);
    const int w = P.col_width[c];
    if (w == 8) gpw<uint64_t>(P.col_out[c])[pos] = v;
    else if (w == 4) gpw<uint32_t>(P.col_out[c])[pos] = (uint32_t)v;
    else gpw<uint8_t>(P.col_out[c])[pos] = (uint8_t)v;
    if (P.col_valid[c]) gpw<uint8_t>(P.col_valid[c])[pos] = ok;
  }
}

// Range and count of the first key words a thread wrote (rows without a NULL
// key), folded per workgroup for the dense layout's choice.
struct KeyRange {
  int64_t lo = INT64_MAX, hi = INT64_MIN;
  int64_t n = 0;
  __device__ __forceinline__ void add(int64_t k) {
    lo = k < lo ? k : lo;
    hi = k > hi ? k : hi;
    n++;
  }
};

// Fold every thread's KeyRange into thread 0's (all threads of the workgroup
// call it; at most 1024 threads).
__device__ __forceinline__ void block_key_range(KeyRange& kr) {
  __shared__ int64_t s_lo[16], s_hi[16], s_n[16];
  for (int o = 32; o > 0; o >>= 1) {
    kr.lo = min(kr.lo, (int64_t)__shfl_down(kr.lo, o));
    kr.hi = max(kr.hi, (int64_t)__shfl_down(kr.hi, o));
    kr.n += (int64_t)__shfl_down(kr.n, o);
  }
  const int wv = threadIdx.x >> 6;
  if (__lane_id() == 0) {
    s_lo[wv] = kr.lo;
    s_hi[wv] = kr.hi;
    s_n[wv] = kr.n;
  }
  __syncthreads();
  if (threadIdx.x == 0)
    for (int w = 1; w < (int)((blockDim.x + 63) >> 6); w++) {
      kr.lo = min(kr.lo, s_lo[w]);
      kr.hi = max(kr.hi, s_hi[w]);
      kr.n += s_n[w];
    }
}

// The workgroup's {min, max, count} at out[0..2] (read by the next kernel).
__device__ __forceinline__ void write_key_range(KeyRange kr, int64_t* out) {
  block_key_range(kr);
  if (threadIdx.x == 0) {
    out[0] = kr.lo;
    out[1] = kr.hi;
    out[2] = kr.n;
  }
}

__device__ __forceinline__ void sink_join_build(const Plan& P, const Code& ins, const uint32_t (&rows)[MAX_STAGES],
                                                KeyRange& kr) {
  const uint64_t pos = P.out_identity ? (uint64_t)rows[0] : wave_alloc(P.out_count, 1);
  if (pos >= P.out_capacity) {
    atomicOr(P.out_overflow, 1u);
    return;
  }
  bool any_null = false;
  uint64_t k0 = 0;
#pragma unroll 1
  for (int k = 0; k < P.nkeys; k++) {
    uint64_t v;
    if (!eval(ins, P.keys[k], rows, v)) {
      any_null = true;
      v = 0;
    }
    if (k == 0) k0 = v;
    P.key_out[pos * P.nkeys + k] = v;
  }
  P.key_null[pos] = any_null;
  if (!any_null) kr.add((int64_t)k0);
  write_cols(P, ins, rows, pos);
}

__device__ __forceinline__ void sink_collect(const Plan& P, const Code& ins, const uint32_t (&rows)[MAX_STAGES]) {
  const uint64_t pos = wave_alloc(P.out_count, 1);
  if (pos >= P.out_capacity) {
    atomicOr(P.out_overflow, 1u);
    return;
  }
  write_cols(P, ins, rows, pos);
  if (P.pair_out) {
    for (int s = 0; s < P.pair_stages; s++) P.pair_out[pos * P.pair_stages + s] = row_of(rows, s);
  }
}

// COLLECT at a position the thread already owns (row-per-lane two-pass collect)
__device__ __forceinline__ void collect_at(const Plan& P, const Code& ins, const uint32_t (&rows)[MAX_STAGES],
                                           uint64_t pos) {
  if (pos >= P.out_capacity) {
    atomicOr(P.out_overflow, 1u);
    return;
  }
  write_cols(P, ins, rows, pos);
  if (P.pair_out) {
    for (int s = 0; s < P.pair_stages; s++) P.pair_out[pos * P.pair_stages + s] = row_of(rows, s);
  }
}

template <int SINK>
__device__ __forceinline__ void do_sink(const Plan& P, const Code& ins, const uint32_t (&rows)[MAX_STAGES], int64_t i,
                                        uint64_t* lds, KeyRange& kr) {
  if constexpr (SINK == SINK_HASH_AGG) sink_hash_agg(P, ins, rows);
  else if constexpr (SINK == SINK_DIRECT_AGG) sink_direct_agg(P, ins, rows, lds);
  else if constexpr (SINK == SINK_JOIN_BUILD) sink_join_build(P, ins, rows, kr);
  else if constexpr (SINK == SINK_COLLECT) sink_collect(P, ins, rows);
  else if constexpr (SINK == SINK_COUNT_PAIRS) wave_alloc(P.out_count, 1);
  else if constexpr (SINK == SINK_PROJECT) write_cols(P, ins, rows, (uint64_t)rows[0]);
  else if constexpr (SINK == SINK_HASH) {
    uint64_t h = HSEED;
#pragma unroll 1
    for (int k = 0; k < P.nkeys; k++) {
      uint64_t v;
      const bool ok = eval(ins, P.keys[k], rows, v);
      const uint64_t x = !ok ? HNULL : (((P.key_prehashed >> k) & 1) ? v : hash_u64(v));
      h = hash_combine(h, x);
    }
    P.hash_out[i] = h;
  }
}

// The fused pipeline: Source -> (filter | probe)* -> sink, one row per lane.
template <int SINK>
__global__ void __launch_bounds__(256) pipeline_kernel(Plan P) {
  extern __shared__ uint64_t lds[];
  const Code ins{(CInstr*)P.ins, P.nins - 1};
  const int ncells = (SINK == SINK_DIRECT_AGG) ? P.ngroups * P.nacc : 0;
  if constexpr (SINK == SINK_DIRECT_AGG) {
    for (int c = threadIdx.x; c < ncells; c += blockDim.x) lds[c] = acc_init_bits(P.acc[c % P.nacc].kind);
    __syncthreads();
  }
  // COUNT_PAIRS: each lane counts its matches in a register and the wave adds
  // its total to the counter once, after the loop (matches in an expand loop
  // are scattered over lanes, and one atomic per match serializes on the counter)
  uint64_t npairs = 0;
  // COLLECT with thread_offsets: this thread's rows go to consecutive positions
  // from its offset (counted by a COUNT_PAIRS pass over the same grid)
  const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t cursor = 0;
  if constexpr (SINK == SINK_COLLECT) cursor = P.thread_offsets ? P.thread_offsets[tid] : 0;
  if constexpr (SINK == SINK_JOIN_BUILD) {
    // every source row is written at its own index: the count is the range
    // (and the overflow word stays clear)
    if (P.out_identity && tid == 0) {
      P.out_count[0] = (uint64_t)(P.end - P.begin);
      P.out_count[1] = 0;
    }
  }
  KeyRange kr;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = P.begin + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < P.end; i += stride) {
    uint32_t rows[MAX_STAGES];
#pragma unroll
    for (int s = 0; s < MAX_STAGES; s++) rows[s] = 0;
    rows[0] = P.sel ? P.sel[i] : (uint32_t)i;
    if (P.expand_step < 0) {
      const bool alive = run_steps(P, ins, 0, P.nsteps, rows);
      if constexpr (SINK == SINK_FLAGS) P.flag_out[i] = alive;
      else if constexpr (SINK == SINK_COUNT_PAIRS) npairs += alive ? 1u : 0u;
      else if constexpr (SINK == SINK_COLLECT) {
        if (alive) {
          if (P.thread_offsets) collect_at(P, ins, rows, cursor++);
          else sink_collect(P, ins, rows);
        }
      } else if (alive) do_sink<SINK>(P, ins, rows, i, lds, kr);
    } else {
      // one probe may match several build rows: the rest of the chain runs per match
      const int e = P.expand_step;
      bool alive = run_steps(P, ins, 0, e, rows);
      if (alive) {
        const Step& st = P.steps[e];
        uint64_t w[MAX_KEYS + 1];
        if (probe_keys(P, ins, st, rows, w)) {
          const DevJoinTable& T = P.tables[st.table];
          const uint64_t h = key_hash(w, st.nkeys);
          uint64_t pos = h & T.mask;
          int64_t row = 0;
          if (bloom_maybe(T, h)) {
            while (join_next(T, w, h, pos, row)) {
              set_row(rows, st.stage, row);
              if (run_steps(P, ins, e + 1, P.nsteps, rows)) {
                if constexpr (SINK == SINK_COUNT_PAIRS) npairs++;
                else if constexpr (SINK == SINK_COLLECT) {
                  if (P.thread_offsets) collect_at(P, ins, rows, cursor++);
                  else sink_collect(P, ins, rows);
                } else do_sink<SINK>(P, ins, rows, i, lds, kr);
              }
              pos = join_step(T, pos);
            }
          }
        }
      }
    }
  }
  if constexpr (SINK == SINK_DIRECT_AGG) {
    __syncthreads();
    uint64_t* out = (uint64_t*)P.direct_partials + (uint64_t)blockIdx.x * ncells;
    for (int c = threadIdx.x; c < ncells; c += blockDim.x) out[c] = lds[c];
  }
  if constexpr (SINK == SINK_JOIN_BUILD) {
    if (P.key_partials) write_key_range(kr, P.key_partials + 3 * (int64_t)blockIdx.x);
  }
  if constexpr (SINK == SINK_COUNT_PAIRS) {
    // per-thread counts for a two-pass collect (a count past 32 bits flags it)
    if (P.wave_counts) {
      if (npairs > 0xFFFFFFFFull) atomicOr(P.out_overflow, 2u);
      P.wave_counts[tid] = (uint32_t)npairs;
    }
    // every lane of the wave is here (the loop above has no early exit)
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) npairs += __shfl_down(npairs, off);
    if (__lane_id() == 0 && npairs) atomicAdd((unsigned long long*)P.out_count, (unsigned long long)npairs);
  }
}

// ------------------------------------------------------ wave-queued pipeline
// Queue heads and counts of levels 1..MAX_LEVELS, 10 bits each (0..1023),
// packed into one wave-uniform 64-bit word so they stay in scalar registers.
__device__ __forceinline__ uint32_t qget(uint64_t packed, int l) { return (uint32_t)(packed >> (10 * (l - 1))) & 1023u; }
__device__ __forceinline__ uint64_t qset(uint64_t packed, int l, uint32_t v) {
  return (packed & ~(1023ull << (10 * (l - 1)))) | ((uint64_t)v << (10 * (l - 1)));
}

// Streaming first probe for FAST_U rows per lane (rows next + u*64 + lane):
// every row's key load, Bloom word and first slot are requested before any
// is used, so a lane keeps FAST_U independent loads in flight per step.
// Streaming key columns are read once: non-temporal loads keep them from
// evicting the Bloom filter and table lines that later rows hit in L2.
#ifndef ML_NT_KEYS
#define ML_NT_KEYS 1
#endif
template <class T>
__device__ __forceinline__ T stream_load(const void* col, uint32_t row) {
#if ML_NT_KEYS
  return __builtin_nontemporal_load(gp<T>(col) + row);
#else
  return gp<T>(col)[row];
#endif
}
__device__ __forceinline__ uint64_t fast_key(uint16_t op, const void* col, uint32_t row) {
  if (op == OP_LOAD_I32) return sx32((uint32_t)stream_load<int32_t>(col, row));
  if (op == OP_LOAD_U32) return (uint64_t)stream_load<uint32_t>(col, row);
  if (op == OP_LOAD_I64) return stream_load<uint64_t>(col, row);
  return stream_load<uint8_t>(col, row) != 0;
}

// FAST_U loads of one column, all issued before any is used: the type
// dispatch is one uniform branch around the unrolled loads (a branch between
// loads would make the compiler wait for the earlier ones at the join).
__device__ __forceinline__ void load_rows(uint16_t op, const void* col, const uint32_t (&row)[FAST_U],
                                          uint64_t (&out)[FAST_U]) {
  if (op == OP_LOAD_I64) {
#pragma unroll
    for (int u = 0; u < FAST_U; u++) out[u] = stream_load<uint64_t>(col, row[u]);
  } else if (op == OP_LOAD_I32) {
#pragma unroll
    for (int u = 0; u < FAST_U; u++) out[u] = sx32((uint32_t)stream_load<int32_t>(col, row[u]));
  } else if (op == OP_LOAD_U32) {
#pragma unroll
    for (int u = 0; u < FAST_U; u++) out[u] = (uint64_t)stream_load<uint32_t>(col, row[u]);
  } else {
#pragma unroll
    for (int u = 0; u < FAST_U; u++) out[u] = stream_load<uint8_t>(col, row[u]) != 0;
  }
}

// Source rows next + u*64 + lane of a level-0 round; rows past `stop` are
// inactive and read the round's first row instead, so every load below is
// unconditional and in bounds.
__device__ __forceinline__ void round_rows(const Plan& P, int64_t next, int64_t stop, int lane,
                                           uint32_t (&src)[FAST_U], bool (&act)[FAST_U]) {
  int64_t idx[FAST_U];
#pragma unroll
  for (int u = 0; u < FAST_U; u++) {
    const int64_t i = next + u * 64 + lane;
    act[u] = i < stop;
    idx[u] = act[u] ? i : next;
  }
  if (P.sel) {
#pragma unroll
    for (int u = 0; u < FAST_U; u++) src[u] = gp<uint32_t>(P.sel)[idx[u]];
  } else {
#pragma unroll
    for (int u = 0; u < FAST_U; u++) src[u] = (uint32_t)idx[u];
  }
}

// Streaming first probe for FAST_U rows per lane: every row's key load, Bloom
// word and slot are requested before any is used, so a lane keeps FAST_U
// independent loads in flight per step.
// The key words of a streaming-probe round (source rows next + u*64 + lane).
struct FastKeys {
  uint32_t src[FAST_U];
  bool act[FAST_U];
  uint64_t k0[FAST_U], k1[FAST_U];
};

__device__ __forceinline__ void fast_keys(const Plan& P, int64_t next, int64_t stop, int lane, FastKeys& K) {
  round_rows(P, next, stop, lane, K.src, K.act);
  load_rows(P.fast_op[0], P.fast_col[0], K.src, K.k0);
  if (P.fast_nkeys > 1) load_rows(P.fast_op[1], P.fast_col[1], K.src, K.k1);
}

__device__ __forceinline__ void fast_probe(const Plan& P, const FastKeys& K, uint32_t (&brow)[FAST_U],
                                           bool (&hit)[FAST_U]) {
  const DevJoinTable& T = P.tables[P.steps[0].table];
  const auto& k0 = K.k0;
  const auto& k1 = K.k1;
  uint64_t h[FAST_U];
#pragma unroll
  for (int u = 0; u < FAST_U; u++) hit[u] = K.act[u];
  const bool two = P.fast_nkeys > 1;
  if (T.dense) {  // (Bloom words,) presence bits of all rows, then the ranks of the present keys' words
    if (T.bloom) {
      uint64_t word[FAST_U], bits[FAST_U];
#pragma unroll
      for (int u = 0; u < FAST_U; u++) {
        uint64_t hh = tmix(k0[u] ^ TSEED);
        if (two) hh = tmix(hh ^ k1[u]);
        bits[u] = bloom_bits(bloom_mix(hh));
        const uint64_t i = k0[u] - (uint64_t)T.dense_min;
        hit[u] = hit[u] && i < T.dense_range;
        word[u] = gp<uint64_t>(T.bloom)[hit[u] ? (i >> DENSE_BLOOM_SHIFT) : 0];
      }
#pragma unroll
      for (int u = 0; u < FAST_U; u++) hit[u] = hit[u] && (word[u] & bits[u]) == bits[u];
    }
    uint64_t idx[FAST_U];
#pragma unroll
    for (int u = 0; u < FAST_U; u++) {
      idx[u] = k0[u] - (uint64_t)T.dense_min;
      hit[u] = hit[u] && idx[u] < T.dense_range;
    }
    if (T.dense_full) {  // every key of the range present: its row is its offset
#pragma unroll
      for (int u = 0; u < FAST_U; u++) brow[u] = (uint32_t)idx[u];
    } else {
      uint64_t bw[FAST_U];
#pragma unroll
      for (int u = 0; u < FAST_U; u++) bw[u] = gp<uint64_t>(T.dense_bits)[hit[u] ? (idx[u] >> 6) : 0];
      if (T.dense_rank) {
        uint32_t rk[FAST_U];
#pragma unroll
        for (int u = 0; u < FAST_U; u++) {
          hit[u] = hit[u] && ((bw[u] >> (idx[u] & 63)) & 1ull);
          rk[u] = gp<uint32_t>(T.dense_rank)[hit[u] ? (idx[u] >> 6) : 0];
        }
#pragma unroll
        for (int u = 0; u < FAST_U; u++)
          brow[u] = rk[u] + (uint32_t)__popcll(bw[u] & ((1ull << (idx[u] & 63)) - 1ull));
      } else {  // rows at their key offset: the bit decides, no rank load
#pragma unroll
        for (int u = 0; u < FAST_U; u++) {
          hit[u] = hit[u] && ((bw[u] >> (idx[u] & 63)) & 1ull);
          brow[u] = (uint32_t)idx[u];
        }
      }
    }
    if (two) {  // the second key word must match the row's
      uint64_t kb[FAST_U];
      if (T.dense_keys8) {
#pragma unroll
        for (int u = 0; u < FAST_U; u++) kb[u] = (uint64_t)gp<uint8_t>(T.dense_keys)[hit[u] ? brow[u] : 0];
      } else if (T.dense_keys32) {
#pragma unroll
        for (int u = 0; u < FAST_U; u++) {
          const uint32_t v = gp<uint32_t>(T.dense_keys)[hit[u] ? brow[u] : 0];
          kb[u] = ((T.dense_keys_sx >> 1) & 1u) ? (uint64_t)(int64_t)(int32_t)v : (uint64_t)v;
        }
      } else {
#pragma unroll
        for (int u = 0; u < FAST_U; u++) kb[u] = gp<uint64_t>(T.dense_keys)[hit[u] ? brow[u] : 0];
      }
#pragma unroll
      for (int u = 0; u < FAST_U; u++) hit[u] = hit[u] && kb[u] == k1[u];
    }
    return;
  }
#pragma unroll
  for (int u = 0; u < FAST_U; u++) {
    uint64_t hh = tmix(k0[u] ^ TSEED);
    if (two) hh = tmix(hh ^ k1[u]);
    h[u] = hh;
  }
  if (T.bloom) {
    uint64_t word[FAST_U], bits[FAST_U];
#pragma unroll
    for (int u = 0; u < FAST_U; u++) {
      const uint64_t g = bloom_mix(h[u]);
      bits[u] = bloom_bits(g);
      word[u] = gp<uint64_t>(T.bloom)[hit[u] ? bloom_word(T, h[u], g) : 0];
    }
#pragma unroll
    for (int u = 0; u < FAST_U; u++) hit[u] = hit[u] && (word[u] & bits[u]) == bits[u];
  }
  // Linear probe of all FAST_U rows in lockstep: each step loads the slot of
  // every row (slot 0 for resolved ones) before testing any of them.
  const int stride = 1 + T.nkw;
  const auto slots = gp<uint64_t>(T.slots);
  uint64_t pos[FAST_U];
  bool pend[FAST_U];
#pragma unroll
  for (int u = 0; u < FAST_U; u++) {
    pos[u] = h[u] & T.mask;
    pend[u] = hit[u];
    hit[u] = false;
    brow[u] = 0;
  }
  for (;;) {
    bool any = false;
#pragma unroll
    for (int u = 0; u < FAST_U; u++) any = any || pend[u];
    if (!__any(any)) break;
    uint64_t hdr[FAST_U], s0[FAST_U], s1[FAST_U];
#pragma unroll
    for (int u = 0; u < FAST_U; u++) {
      const uint64_t at = (pend[u] ? pos[u] : 0) * stride;
      hdr[u] = slots[at];
      s0[u] = slots[at + 1];
    }
    if (two) {
#pragma unroll
      for (int u = 0; u < FAST_U; u++) s1[u] = slots[(pend[u] ? pos[u] : 0) * stride + 2];
    }
#pragma unroll
    for (int u = 0; u < FAST_U; u++) {
      const bool empty = hdr[u] == 0;
      const bool match = !empty && (uint32_t)(hdr[u] >> 32) == (uint32_t)(h[u] >> 32) && s0[u] == k0[u] &&
                         (!two || s1[u] == k1[u]);
      if (pend[u] && match) {
        hit[u] = true;
        brow[u] = (uint32_t)hdr[u] - 1u;
      }
      pend[u] = pend[u] && !empty && !match;
      pos[u] = join_step(T, pos[u]);
    }
  }
}

// Outcome bits of the filter atoms for FAST_U rows: each atom's column is
// loaded for all rows before any comparison.
__device__ __forceinline__ void eval_atoms(const Plan& P, const uint32_t (&src)[FAST_U], uint32_t (&bits)[FAST_U]) {
#pragma unroll
  for (int u = 0; u < FAST_U; u++) bits[u] = 0;
#pragma unroll
  for (int a = 0; a < MAX_ATOMS; a++) {
    if (a >= P.natoms) break;
    const FilterAtom& A = P.atoms[a];
    uint64_t x[FAST_U];
    load_rows((uint16_t)column_load_op(A.op, A.flags), A.col, src, x);
    if (A.op == OP_LOOKUPC) {
#pragma unroll
      for (int u = 0; u < FAST_U; u++) bits[u] |= (uint32_t)(gp<uint8_t>((const void*)A.imm2)[x[u]] != 0) << a;
    } else {
      const uint32_t c = (A.flags >> 8) & 7u;
      const bool f = (A.flags & 0x800u) != 0;
#pragma unroll
      for (int u = 0; u < FAST_U; u++) {
        bool lt, gt, eq;
        if (f) {
          const double p = u2f(x[u]), q = u2f(A.imm2);
          lt = p < q, gt = p > q, eq = p == q;
        } else {
          const int64_t p = (int64_t)x[u], q = (int64_t)A.imm2;
          lt = p < q, gt = p > q, eq = p == q;
        }
        const bool r = c == 0 ? eq : c == 1 ? !eq : c == 2 ? lt : c == 3 ? (lt || eq) : c == 4 ? gt : (gt || eq);
        bits[u] |= (uint32_t)r << a;
      }
    }
  }
}

// Row indices of the stages in `mask`, stored as u32 at base[k * stride]
// for the k-th stage of the mask. Unrolled over stages so rows[] stays in
// registers.
__device__ __forceinline__ void store_stages(uint32_t* base, uint64_t stride, uint32_t mask,
                                             const uint32_t (&rows)[MAX_STAGES]) {
  int k = 0;
#pragma unroll
  for (int s = 0; s < MAX_STAGES; s++)
    if (mask >> s & 1) base[(uint64_t)(k++) * stride] = rows[s];
}
__device__ __forceinline__ void load_stages(const uint32_t* base, uint64_t stride, uint32_t mask,
                                            uint32_t (&rows)[MAX_STAGES]) {
  int k = 0;
#pragma unroll
  for (int s = 0; s < MAX_STAGES; s++)
    if (mask >> s & 1) rows[s] = base[(uint64_t)(k++) * stride];
}

// Each wave owns source rows [begin + w*chunk, begin + (w+1)*chunk). Level 0
// reads 64 source rows; level l >= 1 pops up to 64 queued rows. A round runs
// the level's segment; survivors go to the next level's queue (or the sink
// after the last segment). The deepest level holding a full wave runs first,
// so queues never exceed QCAP entries, and rows reach the sink in source
// order: every queued row precedes the rows of shallower levels. Level 0
// takes FAST_U * 64 rows per round (so level 1 holds up to QCAP_FAST).
#ifndef ML_FLOW_WAVES
#define ML_FLOW_WAVES 0  // minimum waves per SIMD asked of the register allocator (0 = none)
#endif
#ifndef ML_PF_COLS
#define ML_PF_COLS 2     // stage-0 columns loaded ahead at level 0
#endif
#ifndef ML_FAST_PREFETCH
#define ML_FAST_PREFETCH 0  // streaming probe: next round's keys requested before this round's probe
#endif
#ifndef ML_FAST_FLOW_WAVES
#define ML_FAST_FLOW_WAVES 0  // waves per SIMD asked for the streaming-probe form (0 = none; 5 measured no faster)
#endif
// MODE: FLOW_GENERIC (level 0 is a filter of column atoms or the interpreter)
// or FLOW_FAST (level 0 is the streaming first probe). Each form carries only
// its own level-0 code, so it needs fewer registers.
enum : int { FLOW_GENERIC = 0, FLOW_FAST = 1 };
template <int MODE>
constexpr int flow_waves() {
  return MODE == FLOW_FAST ? ML_FAST_FLOW_WAVES : ML_FLOW_WAVES;
}

template <int SINK, int MODE>
__global__ void __launch_bounds__(256, flow_waves<MODE>()) flow_kernel(Plan P) {
  extern __shared__ uint64_t lds[];
  const Code ins{(CInstr*)P.ins, P.nins - 1};
  // direct aggregate: direct_copies copies of the cells, direct_stride words
  // apart (odd, so lanes hit distinct LDS banks); lane l adds to copy
  // l mod copies, so the lanes of a wave rarely add to the same word
  const int ncells = (SINK == SINK_DIRECT_AGG) ? P.ngroups * P.nacc : 0;
  const int cstride = (SINK == SINK_DIRECT_AGG) ? P.direct_stride : 0;
  const int ncopies = (SINK == SINK_DIRECT_AGG) ? P.direct_copies : 0;
  if constexpr (SINK == SINK_DIRECT_AGG) {
    for (int c = threadIdx.x; c < ncopies * cstride; c += blockDim.x) {
      const int cc = c % cstride;
      lds[c] = cc < ncells ? acc_init_bits(P.acc[cc % P.nacc].kind) : 0;
    }
    __syncthreads();
  }
  const int lane = __lane_id();
  const int wv = threadIdx.x >> 6;
  uint64_t* cells = lds + (SINK == SINK_DIRECT_AGG ? (lane & (ncopies - 1)) * cstride : 0);
  uint32_t* q = (uint32_t*)(lds + ncopies * cstride) + wv * P.qwords;
  uint64_t* ac = nullptr;
  if constexpr (SINK == SINK_HASH_AGG) {
    if (P.ac_n) {
      ac = lds + P.ac_off;
      uint32_t* tags = (uint32_t*)(ac + (uint64_t)P.ac_n * P.nacc);
      for (int i = threadIdx.x; i < P.ac_n * P.nacc; i += blockDim.x) ac[i] = acc_init_bits(P.acc[i / P.ac_n].kind);
      for (int i = threadIdx.x; i < P.ac_n; i += blockDim.x) tags[i] = 0;
      __syncthreads();
    }
  }
  const int64_t gw = (int64_t)blockIdx.x * (blockDim.x >> 6) + wv;
  int64_t next = P.begin + gw * P.chunk;
  const int64_t stop = min(P.end, next + P.chunk);
  static_assert(QCAP_FAST <= 1024 && 10 * MAX_LEVELS <= 64, "queue state packing");
  uint64_t qh = 0, qn = 0;
  const int last = P.nseg - 1;
  uint64_t nout = 0;
#if ML_FAST_PREFETCH
  FastKeys fk;
  if (MODE == FLOW_FAST && next < stop) fast_keys(P, next, stop, lane, fk);
#endif
  for (;;) {
    int lvl = -1;
#pragma unroll
    for (int l = MAX_LEVELS; l >= 1; l--)
      if (lvl < 0 && l <= last && qget(qn, l) >= 64) lvl = l;
    if (lvl < 0 && next < stop) lvl = 0;
    if (lvl < 0) {
#pragma unroll
      for (int l = MAX_LEVELS; l >= 1; l--)
        if (lvl < 0 && l <= last && qget(qn, l) > 0) lvl = l;
    }
    if (lvl < 0) break;
    lvl = uni(lvl);

    // survivors of a segment: into the next level's queue, in lane order
    auto push = [&](int L, bool alive, const uint32_t (&rows)[MAX_STAGES]) {
      const uint64_t m = __ballot(alive);
      const uint32_t h = qget(qh, L), c = qget(qn, L);
      if (alive) {
        const uint32_t cap = (uint32_t)P.qcap[L];
        const uint32_t slot = (h + c + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))) & (cap - 1);
        store_stages(q + P.qoff[L] + slot, cap, P.live[L], rows);
      }
      qn = qset(qn, L, c + (uint32_t)__popcll(m));
    };
    // rows that passed the last segment
    auto sink = [&](bool alive, const uint32_t (&rows)[MAX_STAGES]) {
      if constexpr (SINK == SINK_DIRECT_AGG) {
        if (alive) sink_direct_agg(P, ins, rows, cells);
      } else if constexpr (SINK == SINK_HASH_AGG) {
        if (alive) {
          if (ac) sink_hash_agg_cached(P, ins, rows, ac);
          else sink_hash_agg(P, ins, rows);
        }
      } else {
        const uint64_t m = __ballot(alive);
        if constexpr (SINK == SINK_ROWS) {
          if (alive) {
            const uint64_t pos = (uint64_t)gw * P.chunk + nout + (uint64_t)__popcll(m & ((1ull << lane) - 1ull));
            store_stages(P.tuples + pos, P.staging, P.tup_mask, rows);
          }
        }
        nout += (uint64_t)__popcll(m);
      }
    };

    if (MODE == FLOW_FAST && lvl == 0) {
      // FAST_U * 64 source rows through the streaming probe; hits to level 1 in source order
      uint32_t brow[FAST_U];
      bool hit[FAST_U];
#if ML_FAST_PREFETCH == 1
      // this round's keys were requested by the previous level-0 round (or
      // before the loop); request the next round's before probing this one
      const FastKeys K = fk;
      next = min(stop, next + FAST_U * 64);
      if (next < stop) fast_keys(P, next, stop, lane, fk);
#elif ML_FAST_PREFETCH == 2
      // this round's keys were requested at the end of the previous level-0
      // round (or before the loop); the next round's are requested below,
      // after this round's table loads have returned
      const FastKeys K = fk;
      next = min(stop, next + FAST_U * 64);
#else
      FastKeys K;
      fast_keys(P, next, stop, lane, K);
      next = min(stop, next + FAST_U * 64);
#endif
      fast_probe(P, K, brow, hit);
      const int stage = P.steps[0].stage;
#pragma unroll
      for (int u = 0; u < FAST_U; u++) {
        uint32_t rows[MAX_STAGES];
#pragma unroll
        for (int s2 = 0; s2 < MAX_STAGES; s2++) rows[s2] = 0;
        rows[0] = K.src[u];
        set_row(rows, stage, brow[u]);
        push(1, hit[u], rows);
      }
#if ML_FAST_PREFETCH == 2
      // vector loads return in issue order: keys requested before the table
      // loads would make every table wait for them, requested here they stay
      // in flight through the deeper levels' rounds
      __builtin_amdgcn_sched_barrier(0);
      if (next < stop) fast_keys(P, next, stop, lane, fk);
#endif
    } else if (MODE == FLOW_GENERIC && lvl == 0 && P.natoms) {
      // FAST_U * 64 source rows through a filter of column atoms
      uint32_t src[FAST_U], bits[FAST_U];
      bool act[FAST_U];
      round_rows(P, next, stop, lane, src, act);
      bool batched = false;
      if constexpr (SINK == SINK_DIRECT_AGG && FAST_U == 4) {
        if (last == 0 && P.src_batch) {
          // the sink evaluates all 4 rows of the lane at once, from columns
          // requested together with the filter's before any is used
          Staged8 T;
          stage_load(P, src, T);
          eval_atoms(P, src, bits);
          bool alive[4];
#pragma unroll
          for (int u = 0; u < 4; u++) alive[u] = act[u] && ((P.atom_truth >> bits[u]) & 1u);
          sink_direct_agg_src<4>(P, ins, src, T, alive, cells);
          batched = true;
        }
      }
      if (!batched) eval_atoms(P, src, bits);
      next = min(stop, next + FAST_U * 64);
      if (!batched) {
#pragma unroll
        for (int u = 0; u < FAST_U; u++) {
          uint32_t rows[MAX_STAGES];
#pragma unroll
          for (int s2 = 0; s2 < MAX_STAGES; s2++) rows[s2] = 0;
          rows[0] = src[u];
          const bool alive = act[u] && ((P.atom_truth >> bits[u]) & 1u);
          if (last == 0) sink(alive, rows);
          else push(1, alive, rows);
        }
      }
    } else if (MODE == FLOW_GENERIC && lvl == 0) {
      // FAST_U * 64 source rows; the stage-0 columns segment 0 reads are
      // loaded for all of them first, then the segment runs row by row
      uint32_t src[FAST_U];
      bool act[FAST_U];
      uint64_t v0[FAST_U], v1[FAST_U];
      round_rows(P, next, stop, lane, src, act);
#pragma unroll
      for (int u = 0; u < FAST_U; u++) v0[u] = v1[u] = 0;
      if (P.pf_n > 0) load_rows(P.pf_op[0], P.pf_col[0], src, v0);
      if (ML_PF_COLS > 1 && P.pf_n > 1) load_rows(P.pf_op[1], P.pf_col[1], src, v1);
      next = min(stop, next + FAST_U * 64);
#pragma unroll 1
      for (int u = 0; u < FAST_U; u++) {
        uint32_t rows[MAX_STAGES];
#pragma unroll
        for (int s2 = 0; s2 < MAX_STAGES; s2++) rows[s2] = 0;
        Pf pf;
        pf.c0 = P.pf_n > 0 ? P.pf_col[0] : nullptr;
        pf.c1 = (ML_PF_COLS > 1 && P.pf_n > 1) ? P.pf_col[1] : nullptr;
        // pick the u-th values with selects (the arrays stay in registers)
#pragma unroll
        for (int k = 0; k < FAST_U; k++) {
          if (k == u) {
            rows[0] = src[k];
            pf.v0 = v0[k];
            pf.v1 = v1[k];
          }
        }
        bool alive = false;
#pragma unroll
        for (int k = 0; k < FAST_U; k++)
          if (k == u) alive = act[k];
        alive = alive && run_steps(P, ins, 0, P.seg_end[0], rows, pf);
        if (last == 0) sink(alive, rows);
        else push(1, alive, rows);
      }
    } else {
      uint32_t rows[MAX_STAGES];
#pragma unroll
      for (int s = 0; s < MAX_STAGES; s++) rows[s] = 0;
      const uint32_t h = qget(qh, lvl), c = qget(qn, lvl);
      const int n = (int)min(64u, c);
      const uint32_t cap = (uint32_t)P.qcap[lvl];
      if (lane < n) load_stages(q + P.qoff[lvl] + ((h + (uint32_t)lane) & (cap - 1)), cap, P.live[lvl], rows);
      qh = qset(qh, lvl, (h + (uint32_t)n) & (cap - 1));
      qn = qset(qn, lvl, c - (uint32_t)n);
      const bool alive = lane < n && run_steps(P, ins, P.seg_end[lvl - 1], P.seg_end[lvl], rows);
      if (lvl == last) sink(alive, rows);
      else push(lvl + 1, alive, rows);
    }
    // queue writes of this round are read by other lanes in later rounds
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  if constexpr (SINK == SINK_ROWS || SINK == SINK_COUNT_PAIRS) {
    if (lane == 0) P.wave_counts[gw] = (uint32_t)nout;
  }
  if constexpr (SINK == SINK_HASH_AGG) {
    if (ac) {  // merge the cached groups into the table
      __syncthreads();
      const uint32_t* tags = (const uint32_t*)(ac + (uint64_t)P.ac_n * P.nacc);
      for (int i = threadIdx.x; i < P.ac_n; i += blockDim.x) {
        const uint32_t t = tags[i];
        if (!t) continue;
        for (int s = 0; s < P.nacc; s++) {
          const uint64_t v = ac[(uint64_t)s * P.ac_n + i];
          if (v != acc_init_bits(P.acc[s].kind)) acc_merge_global(P.acc[s].kind, P.acc_data[s], (int64_t)(t - 1u), v);
        }
      }
    }
  }
  if constexpr (SINK == SINK_DIRECT_AGG) {
    __syncthreads();
    uint64_t* out = (uint64_t*)P.direct_partials + (uint64_t)blockIdx.x * ncells;
    for (int c = threadIdx.x; c < ncells; c += blockDim.x) {
      const int kind = P.acc[c % P.nacc].kind;
      uint64_t v = lds[c];
      for (int k = 1; k < ncopies; k++) v = acc_combine(kind, v, lds[k * cstride + c]);
      out[c] = v;
    }
  }
}

// Materialize the rows a ROWS flow produced: one workgroup per wave region,
// written densely at the region's exclusive offset (source order kept).
template <int SINK>
__global__ void __launch_bounds__(256) materialize_kernel(Plan P, const uint64_t* offsets) {
  const Code ins{(CInstr*)P.ins, P.nins - 1};
  const int64_t w = blockIdx.x;
  const uint32_t cnt = P.wave_counts[w];
  const uint64_t base = offsets[w];
  KeyRange kr;
  for (uint32_t k = threadIdx.x; k < cnt; k += blockDim.x) {
    const uint64_t src = (uint64_t)w * P.chunk + k;
    uint32_t rows[MAX_STAGES];
#pragma unroll
    for (int s = 0; s < MAX_STAGES; s++) rows[s] = 0;
    load_stages(P.tuples + src, P.staging, P.tup_mask, rows);
    const uint64_t pos = base + k;
    if constexpr (SINK == SINK_JOIN_BUILD) {
      bool any_null = false;
      uint64_t k0 = 0;
#pragma unroll 1
      for (int kk = 0; kk < P.nkeys; kk++) {
        uint64_t v;
        if (!eval(ins, P.keys[kk], rows, v)) {
          any_null = true;
          v = 0;
        }
        if (kk == 0) k0 = v;
        P.key_out[pos * P.nkeys + kk] = v;
      }
      P.key_null[pos] = any_null;
      if (!any_null) kr.add((int64_t)k0);
    }
    write_cols(P, ins, rows, pos);
    if constexpr (SINK == SINK_COLLECT) {
      if (P.pair_out)
        for (int s = 0; s < P.pair_stages; s++) P.pair_out[pos * P.pair_stages + s] = row_of(rows, s);
    }
  }
  if constexpr (SINK == SINK_JOIN_BUILD) {
    if (P.key_partials) write_key_range(kr, P.key_partials + 3 * w);
  }
}

// ---------------------------------------------------------- direct agg reduce
// One workgroup per (group, state) cell: a fixed-shape tree over the block
// partials (thread t folds partials t, t+256, ... in order, then a pairwise
// LDS tree), added to the running total of earlier batches. The shape depends
// only on the partial count, so results are identical from run to run.
__global__ void __launch_bounds__(256) direct_tree_kernel(const uint64_t* partials, int nblocks, int ncells, int nacc,
                                                          AccSlotList accs, uint64_t* out) {
  __shared__ uint64_t buf[256];
  const int c = blockIdx.x;
  const int kind = accs.kind[c % nacc];
  uint64_t acc = acc_init_bits(kind);
  for (int b = threadIdx.x; b < nblocks; b += blockDim.x) acc = acc_combine(kind, acc, partials[(uint64_t)b * ncells + c]);
  buf[threadIdx.x] = acc;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) buf[threadIdx.x] = acc_combine(kind, buf[threadIdx.x], buf[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[c] = acc_combine(kind, out[c], buf[0]);
}

// ------------------------------------------------------------ fills / gathers
__global__ void fill_u64_kernel(uint64_t* p, uint64_t v, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

__global__ void gather_kernel(const void* src, int width, const int64_t* idx, int64_t n, void* dst,
                              const uint8_t* src_valid, uint8_t* dst_valid) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const int64_t r = idx[i];
    if (width == 8) ((uint64_t*)dst)[i] = ((const uint64_t*)src)[r];
    else if (width == 4) ((uint32_t*)dst)[i] = ((const uint32_t*)src)[r];
    else ((uint8_t*)dst)[i] = ((const uint8_t*)src)[r];
    if (dst_valid) dst_valid[i] = src_valid ? src_valid[r] : 1;
  }
}

// ------------------------------------------------------------------ join table
__global__ void join_insert_kernel(DevJoinTable T, const uint64_t* keys, const uint8_t* key_null, int64_t n,
                                   uint64_t* bloom) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int nkw = T.nkw;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += stride) {
    if (key_null[r]) continue;  // NULL build keys are never inserted (join.rs:182-184)
    uint64_t w[MAX_KEYS + 1];
#pragma unroll
    for (int k = 0; k < MAX_KEYS + 1; k++) w[k] = k < nkw ? keys[r * nkw + k] : 0;
    const uint64_t h = key_hash(w, nkw);
    const uint64_t hdr = ((h >> 32) << 32) | (uint64_t)(r + 1);
    uint64_t pos = h & T.mask;
    for (;;) {
      uint64_t* s = T.slots + pos * (1 + nkw);
      if (atomicCAS((unsigned long long*)s, 0ull, (unsigned long long)hdr) == 0ull) {
#pragma unroll
        for (int k = 0; k < MAX_KEYS; k++)
          if (k < nkw) s[1 + k] = w[k];
        break;
      }
      pos = join_step(T, pos);
    }
    if (bloom) {
      const uint64_t g = bloom_mix(h);
      atomicOr((unsigned long long*)&bloom[bloom_word(T, h, g)], (unsigned long long)bloom_bits(g));
    }
  }
}

// unique = 0 if two occupied slots hold the same key (walk each slot's run from its home).
__global__ void join_unique_kernel(DevJoinTable T, uint32_t* dup_flag) {
  const uint64_t cap = T.mask + 1;
  const int nkw = T.nkw;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < (int64_t)cap; p += stride) {
    const uint64_t* s = T.slots + (uint64_t)p * (1 + nkw);
    if (s[0] == 0) continue;
    uint64_t w[MAX_KEYS + 1];
#pragma unroll
    for (int k = 0; k < MAX_KEYS + 1; k++) w[k] = k < nkw ? s[1 + k] : 0;
    const uint64_t h = key_hash(w, nkw);
    uint64_t q = h & T.mask;
    while (q != (uint64_t)p) {
      const uint64_t* t = T.slots + q * (1 + nkw);
      bool eq = t[0] != 0 && (t[0] >> 32) == (s[0] >> 32);
#pragma unroll
      for (int k = 0; k < MAX_KEYS; k++)
        if (k < nkw) eq = eq && t[1 + k] == w[k];
      if (eq) {
        atomicOr(dup_flag, 1u);
        break;
      }
      q = join_step(T, q);
    }
  }
}

// ------------------------------------------------- dense join table build
// min / max and count of the non-NULL key words. Every workgroup writes its
// partial triple; the last one to finish (counted in *done) folds them into
// mm[0] = min, mm[1] = max, mm[2] = rows, mm[3] = non-NULL rows, sets *done
// back to 0 for the next launch and clears the two flag words clear[0..1]
// that the dense build uses next.
__global__ void __launch_bounds__(256) key_minmax_kernel(const uint64_t* keys, int nkw, const uint8_t* key_null,
                                                         int64_t n, const uint64_t* n_dev, int64_t* partials,
                                                         int64_t* mm, uint32_t* done, uint32_t* clear) {
  __shared__ int64_t lo[256], hi[256], cnt[256];
  __shared__ int last;
  if (n_dev) n = min(n, (int64_t)*n_dev);
  int64_t a = INT64_MAX, b = INT64_MIN, c = 0;
  // four independent rows per thread per step (loads issued before use)
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t r0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r0 < n; r0 += 4 * stride) {
    int64_t k[4];
    bool ok[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int64_t r = r0 + u * stride;
      const int64_t rr = r < n ? r : r0;
      k[u] = (int64_t)keys[rr * nkw];
      ok[u] = r < n && !key_null[rr];
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      if (!ok[u]) continue;
      a = k[u] < a ? k[u] : a;
      b = k[u] > b ? k[u] : b;
      c++;
    }
  }
  lo[threadIdx.x] = a;
  hi[threadIdx.x] = b;
  cnt[threadIdx.x] = c;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      lo[threadIdx.x] = min(lo[threadIdx.x], lo[threadIdx.x + o]);
      hi[threadIdx.x] = max(hi[threadIdx.x], hi[threadIdx.x + o]);
      cnt[threadIdx.x] += cnt[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    // agent-scope stores, completed before the count that publishes them
    __hip_atomic_store(partials + 3 * blockIdx.x, lo[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(partials + 3 * blockIdx.x + 1, hi[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(partials + 3 * blockIdx.x + 2, cnt[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    wait_for_stores();
    last = atomicAdd(done, 1u) == gridDim.x - 1;
  }
  __syncthreads();
  if (!last) return;
  // the partials are read with agent-scope loads below
  a = INT64_MAX;
  b = INT64_MIN;
  c = 0;
  for (int i = threadIdx.x; i < (int)gridDim.x; i += blockDim.x) {
    a = min(a, __hip_atomic_load(partials + 3 * i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    b = max(b, __hip_atomic_load(partials + 3 * i + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    c += __hip_atomic_load(partials + 3 * i + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  lo[threadIdx.x] = a;
  hi[threadIdx.x] = b;
  cnt[threadIdx.x] = c;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      lo[threadIdx.x] = min(lo[threadIdx.x], lo[threadIdx.x + o]);
      hi[threadIdx.x] = max(hi[threadIdx.x], hi[threadIdx.x + o]);
      cnt[threadIdx.x] += cnt[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    mm[0] = lo[0];
    mm[1] = hi[0];
    mm[2] = n;
    mm[3] = cnt[0];
    *done = 0;
    clear[0] = 0;
    clear[1] = 0;
  }
}

// Presence bits of the non-NULL keys (key k -> bit k - kmin); flag bit 0: a
// key seen twice, bit 2: a key outside [kmin, kmin + range) (not marked). Lanes whose keys fall in the same
// directory word (runs of neighbouring lanes, as sorted keys give) OR their
// bits together first, so each run costs one atomic.
__global__ void dense_mark_kernel(const uint64_t* keys, int nkw, const uint8_t* key_null, int64_t n, int64_t kmin,
                                  uint64_t range, uint64_t* bits, uint32_t* flag) {
  const int lane = __lane_id();
  bool dup = false, outside = false, wide = false;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63); base < n; base += stride) {
    const int64_t r = base + lane;
    uint64_t w = ~0ull, v = 0;
    if (r < n && !key_null[r]) {
      if (nkw == 2) wide = wide || keys[r * 2 + 1] > 255u;
      const uint64_t i = (uint64_t)((int64_t)keys[r * nkw] - kmin);
      if (i < range) {
        w = i >> 6;
        v = 1ull << (i & 63);
      } else {
        outside = true;
      }
    }
    // segmented suffix OR over runs of equal words: afterwards the first lane
    // of a run holds the run's bits
    const uint64_t wn = __shfl_down(w, 1);
    int end = (lane == 63 || wn != w) ? 1 : 0;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint64_t vn = __shfl_down(v, d);
      const int en = __shfl_down(end, d);
      if (!end && lane + d < 64) {
        dup = dup || (v & vn) != 0;
        v |= vn;
        end = en;
      }
    }
    const uint64_t wp = __shfl_up(w, 1);
    if ((lane == 0 || wp != w) && w != ~0ull) {
      const unsigned long long old = atomicOr((unsigned long long*)&bits[w], (unsigned long long)v);
      dup = dup || (old & v) != 0;
    }
  }
  if (__any(dup) && lane == 0) atomicOr(flag, 1u);
  if (__any(outside) && lane == 0) atomicOr(flag, 4u);
  if (__any(wide) && lane == 0) atomicOr(flag, 256u);
}

// Build row r's key words 1.. and columns, written at table position pos.
__device__ __forceinline__ void dense_write_row(const DensePermute& M, const uint64_t* keys, int64_t r, uint64_t pos) {
  // (a column with dst null is not stored: the table's first-key column)
  for (int k = 1; k < M.nkw; k++) {
    const uint64_t at = pos * (M.nkw - 1) + k - 1;
    if (M.keys32) ((uint32_t*)M.dst_keys)[at] = (uint32_t)keys[r * M.nkw + k];
    else M.dst_keys[at] = keys[r * M.nkw + k];
  }
  if (M.dst_keys8) M.dst_keys8[pos] = (uint8_t)keys[r * 2 + 1];
  for (int c = 0; c < M.ncols; c++) {
    if (!M.dst[c]) continue;
    const int w = M.width[c];
    if (w == 8) ((uint64_t*)M.dst[c])[pos] = ((const uint64_t*)M.src[c])[r];
    else if (w == 4) ((uint32_t*)M.dst[c])[pos] = ((const uint32_t*)M.src[c])[r];
    else ((uint8_t*)M.dst[c])[pos] = ((const uint8_t*)M.src[c])[r];
    if (M.src_valid[c]) M.dst_valid[c][pos] = M.src_valid[c][r];
  }
}

// Build rows into key order: a row with key k moves to the rank of k; rows
// with a NULL key (never matched) follow the present keys in any order.
__global__ void dense_permute_kernel(DensePermute M, const uint64_t* keys, const uint8_t* key_null, int64_t n,
                                     int64_t kmin, uint64_t range, const uint64_t* bits, const uint32_t* rank,
                                     const uint32_t* present, uint32_t* null_next) {
  // rank == null: rows sit at their key's offset (a full table, or an
  // offset-indexed one) and the NULL-key rows start at range
  const uint64_t null_base = rank ? (uint64_t)*present : range;
  const uint64_t held = M.held > 0 ? (uint64_t)M.held : (uint64_t)n;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
    uint64_t pos;
    if (!key_null[r]) {
      const uint64_t i = (uint64_t)((int64_t)keys[r * M.nkw] - kmin);
      if (i >= range) continue;  // flagged by the marking; the table falls back
      pos = rank ? rank[i >> 6] + (uint64_t)__popcll(bits[i >> 6] & ((1ull << (i & 63)) - 1ull)) : i;
      if (pos >= held) continue;  // only with repeated keys, which also fall back
    } else {
      pos = null_base + atomicAdd(null_next, 1u);
      if (pos >= held) continue;  // only with repeated keys, which also fall back
    }
    dense_write_row(M, keys, r, pos);
  }
}

// A table whose rows sit at their key's offset (a full table, or an
// offset-indexed one) in one pass: each row sets its key's presence bit, with
// the flags of dense_mark_kernel, and moves to that offset; NULL-key rows
// follow from `range` on, in any order (*null_next from 0). A repeated key
// leaves overwritten rows behind, and its flag sends the table to the hashed
// layout.
__global__ void dense_place_kernel(DensePermute M, const uint64_t* keys, const uint8_t* key_null, int64_t n,
                                   int64_t kmin, uint64_t range, uint64_t* bits, uint32_t* flag, uint32_t* null_next) {
  const int lane = __lane_id();
  const int nkw = M.nkw;
  const uint64_t held = M.held > 0 ? (uint64_t)M.held : (uint64_t)n;
  bool dup = false, outside = false, wide = false;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  // every lane of a wave takes the same trips (the shuffles below need all 64)
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63); base < n; base += stride) {
    const int64_t r = base + lane;
    uint64_t w = ~0ull, v = 0, pos = ~0ull;
    const bool live = r < n;
    const bool is_null = live && key_null[r];
    if (live && !is_null) {
      if (nkw == 2) wide = wide || keys[r * 2 + 1] > 255u;
      const uint64_t i = (uint64_t)((int64_t)keys[r * nkw] - kmin);
      if (i < range) {
        w = i >> 6;
        v = 1ull << (i & 63);
        pos = i;
      } else {
        outside = true;
      }
    }
    if (is_null) pos = range + atomicAdd(null_next, 1u);
    // lanes whose keys share a presence word OR their bits first (as in
    // dense_mark_kernel): one atomic per run of a word
    const uint64_t wn = __shfl_down(w, 1);
    int end = (lane == 63 || wn != w) ? 1 : 0;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint64_t vn = __shfl_down(v, d);
      const int en = __shfl_down(end, d);
      if (!end && lane + d < 64) {
        dup = dup || (v & vn) != 0;
        v |= vn;
        end = en;
      }
    }
    const uint64_t wp = __shfl_up(w, 1);
    if ((lane == 0 || wp != w) && w != ~0ull) {
      const unsigned long long old = atomicOr((unsigned long long*)&bits[w], (unsigned long long)v);
      dup = dup || (old & v) != 0;
    }
    if (pos < held) dense_write_row(M, keys, r, pos);  // beyond held only with repeated keys
  }
  if (__any(dup) && lane == 0) atomicOr(flag, 1u);
  if (__any(outside) && lane == 0) atomicOr(flag, 4u);
  if (__any(wide) && lane == 0) atomicOr(flag, 256u);
}

// Fold the per-workgroup {min, max, count} partials a join build's row
// kernel wrote into mm = {min, max, rows, non-NULL rows} (as key_minmax_kernel)
// and clear the two flag words clear[0..1] that the dense build uses next.
__global__ void __launch_bounds__(256) key_range_reduce_kernel(const int64_t* partials, int nparts, int64_t n,
                                                               const uint64_t* n_dev, int64_t* mm, uint32_t* clear) {
  KeyRange kr;
  for (int i = threadIdx.x; i < nparts; i += blockDim.x) {
    kr.lo = min(kr.lo, partials[3 * i]);
    kr.hi = max(kr.hi, partials[3 * i + 1]);
    kr.n += partials[3 * i + 2];
  }
  block_key_range(kr);
  if (threadIdx.x == 0) {
    mm[0] = kr.lo;
    mm[1] = kr.hi;
    mm[2] = n_dev ? min(n, (int64_t)*n_dev) : n;
    mm[3] = kr.n;
    clear[0] = 0;
    clear[1] = 0;
  }
}

// Dense build of a small table in one workgroup (LDS): min/max of the first
// key words, presence bits, ranks and the rows moved into key order, for at
// most `cap` rows (*count_dev of them valid) and a key range of at most
// SMALL_DENSE_WORDS * 64. status = {rows, min, range, flags}: flags bit 0 a
// repeated key, bit 1 not dense (no rows, or the range is too wide); with
// flags != 0 nothing but status is written.
constexpr int SMALL_DENSE_WORDS = 4096;
__global__ void __launch_bounds__(1024) small_dense_kernel(DensePermute M, const uint64_t* keys, const uint8_t* key_null,
                                                           int64_t cap, const uint64_t* count_dev, uint64_t* bits_out,
                                                           uint32_t* rank_out, int64_t* status) {
  __shared__ uint64_t sbits[SMALL_DENSE_WORDS];
  __shared__ uint32_t srank[SMALL_DENSE_WORDS + 1];
  __shared__ int64_t wlo[16], whi[16];
  __shared__ uint32_t wsum[16];
  __shared__ int64_t smin;
  __shared__ uint64_t srange;
  __shared__ uint32_t sflag, snull;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int64_t n = min(cap, (int64_t)*count_dev);
  const int nkw = M.nkw;
  int64_t lo = INT64_MAX, hi = INT64_MIN;
  for (int64_t r = tid; r < n; r += blockDim.x) {
    if (key_null[r]) continue;
    const int64_t k = (int64_t)keys[r * nkw];
    lo = k < lo ? k : lo;
    hi = k > hi ? k : hi;
  }
  for (int o = 32; o > 0; o >>= 1) {
    lo = min(lo, (int64_t)__shfl_down(lo, o));
    hi = max(hi, (int64_t)__shfl_down(hi, o));
  }
  if (lane == 0) {
    wlo[wv] = lo;
    whi[wv] = hi;
  }
  __syncthreads();
  if (tid == 0) {
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) {
      lo = min(lo, wlo[w]);
      hi = max(hi, whi[w]);
    }
    uint32_t f = 2u;
    uint64_t R = 0;
    if (n > 0 && lo <= hi) {
      const unsigned __int128 range = (unsigned __int128)((__int128)hi - (__int128)lo) + 1;
      const unsigned __int128 limit = (unsigned __int128)max<uint64_t>((uint64_t)n * 8, 1ull << 20);
      if (range <= limit && range <= (unsigned __int128)SMALL_DENSE_WORDS * 64) {
        f = 0u;
        R = (uint64_t)range;
      }
    }
    smin = lo;
    srange = R;
    sflag = f;
    snull = 0;
  }
  __syncthreads();
  const int64_t kmin = smin;
  const uint64_t R = srange;
  const int ne = (int)((R + 63) / 64);
  if (sflag == 0) {
    for (int w = tid; w < ne; w += blockDim.x) sbits[w] = 0;
    __syncthreads();
    bool dup = false;
    for (int64_t r = tid; r < n; r += blockDim.x) {
      if (key_null[r]) continue;
      const uint64_t i = (uint64_t)((int64_t)keys[r * nkw] - kmin);
      const unsigned long long bit = 1ull << (i & 63);
      dup = dup || (atomicOr((unsigned long long*)&sbits[i >> 6], bit) & bit) != 0;
    }
    if (dup) atomicOr(&sflag, 1u);
    __syncthreads();
  }
  if (sflag != 0) {
    if (tid == 0) {
      status[0] = n;
      status[1] = kmin;
      status[2] = (int64_t)R;
      status[3] = sflag;
    }
    return;
  }
  // ranks: thread t sums the popcounts of words 4t .. 4t+3, then a block scan
  uint32_t c[4], local = 0;
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int w = 4 * tid + j;
    c[j] = w < ne ? (uint32_t)__popcll(sbits[w]) : 0u;
    local += c[j];
  }
  uint32_t incl = local;
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t t = __shfl_up(incl, o);
    if (lane >= o) incl += t;
  }
  if (lane == 63) wsum[wv] = incl;
  __syncthreads();
  uint32_t base = 0, total = 0;
  for (int w = 0; w < (int)(blockDim.x >> 6); w++) {
    if (w < wv) base += wsum[w];
    total += wsum[w];
  }
  uint32_t run = base + incl - local;
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int w = 4 * tid + j;
    if (w < ne) srank[w] = run;
    run += c[j];
  }
  if (tid == 0) srank[ne] = total;
  __syncthreads();
  for (int w = tid; w <= ne; w += blockDim.x) {
    if (w < ne) bits_out[w] = sbits[w];
    rank_out[w] = srank[w];
  }
  for (int64_t r = tid; r < n; r += blockDim.x) {
    uint64_t pos;
    if (!key_null[r]) {
      const uint64_t i = (uint64_t)((int64_t)keys[r * nkw] - kmin);
      pos = srank[i >> 6] + (uint64_t)__popcll(sbits[i >> 6] & ((1ull << (i & 63)) - 1ull));
    } else {
      pos = (uint64_t)total + atomicAdd(&snull, 1u);
    }
    for (int k = 1; k < nkw; k++) {
      const uint64_t at = pos * (nkw - 1) + k - 1;
      if (M.keys32) ((uint32_t*)M.dst_keys)[at] = (uint32_t)keys[r * nkw + k];
      else M.dst_keys[at] = keys[r * nkw + k];
    }
    for (int col = 0; col < M.ncols; col++) {
      const int w = M.width[col];
      if (w == 8) ((uint64_t*)M.dst[col])[pos] = ((const uint64_t*)M.src[col])[r];
      else if (w == 4) ((uint32_t*)M.dst[col])[pos] = ((const uint32_t*)M.src[col])[r];
      else ((uint8_t*)M.dst[col])[pos] = ((const uint8_t*)M.src[col])[r];
      if (M.src_valid[col]) M.dst_valid[col][pos] = M.src_valid[col][r];
    }
  }
  if (tid == 0) {
    status[0] = n;
    status[1] = kmin;
    status[2] = (int64_t)R;
    status[3] = (uint64_t)total == R ? SMALL_DENSE_FULL : 0;
  }
}

// Bloom bits of every non-NULL row's whole key in the word of its first key
// (see DevJoinTable::dense): lanes with the same word (runs of neighbouring
// lanes, as rows in key order give) OR their bits first, one atomic per run.
__global__ void dense_bloom_kernel(DevJoinTable T, const uint64_t* keys, const uint8_t* key_null, int64_t n,
                                   int64_t kmin, uint64_t range, uint64_t* bloom) {
  const int lane = __lane_id();
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63); base < n; base += stride) {
    const int64_t r = base + lane;
    uint64_t w = ~0ull, v = 0;
    if (r < n && !key_null[r]) {
      uint64_t kw[MAX_KEYS + 1];
#pragma unroll
      for (int k = 0; k < MAX_KEYS + 1; k++) kw[k] = k < T.nkw ? keys[r * T.nkw + k] : 0;
      const uint64_t i = (uint64_t)((int64_t)kw[0] - kmin);
      if (i < range) {
        w = i >> DENSE_BLOOM_SHIFT;
        v = bloom_bits(bloom_mix(key_hash(kw, T.nkw)));
      }
    }
    const uint64_t wn = __shfl_down(w, 1);
    int end = (lane == 63 || wn != w) ? 1 : 0;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint64_t vn = __shfl_down(v, d);
      const int en = __shfl_down(end, d);
      if (!end && lane + d < 64) {
        v |= vn;
        end = en;
      }
    }
    const uint64_t wp = __shfl_up(w, 1);
    if ((lane == 0 || wp != w) && w != ~0ull) atomicOr((unsigned long long*)&bloom[w], (unsigned long long)v);
  }
}

// ------------------------------------------- partitioned join table build
// Keys go to the partition of their home slot: (hash & mask) >> part_shift.
// Workgroup b of the histogram and scatter passes owns build rows
// [b*rows_per_block, (b+1)*rows_per_block); hist is partition-major
// (hist[p * nblocks + b]), so its exclusive scan gives each (partition,
// workgroup) pair its range of the partition-ordered staging arrays.
__device__ __forceinline__ void load_key(const uint64_t* keys, int64_t r, int nkw, uint64_t (&w)[MAX_KEYS + 1]) {
#pragma unroll
  for (int k = 0; k < MAX_KEYS + 1; k++) w[k] = k < nkw ? keys[r * nkw + k] : 0;
}

// PART_U rows per thread per step: all their key words are requested before
// any is hashed. Rows past `hi` read row `lo` and are inactive, so every load
// is unconditional and in bounds; NULL build keys are never inserted
// (join.rs:182-184).
constexpr int PART_U = 4;
__device__ __forceinline__ void load_keys_u(const uint64_t* keys, const uint8_t* key_null, int64_t r0, int64_t lo,
                                            int64_t hi, int nkw, uint64_t (&w)[PART_U][MAX_KEYS + 1],
                                            int64_t (&row)[PART_U], bool (&ok)[PART_U]) {
  uint8_t nul[PART_U];
#pragma unroll
  for (int u = 0; u < PART_U; u++) {
    const int64_t r = r0 + (int64_t)u * blockDim.x;
    row[u] = r < hi ? r : lo;
    ok[u] = r < hi;
    nul[u] = key_null[row[u]];
    load_key(keys, row[u], nkw, w[u]);
  }
#pragma unroll
  for (int u = 0; u < PART_U; u++) ok[u] = ok[u] && !nul[u];
}

__global__ void __launch_bounds__(256) join_part_hist_kernel(DevJoinTable T, const uint64_t* keys,
                                                             const uint8_t* key_null, int64_t n,
                                                             int64_t rows_per_block, int nparts, uint32_t* hist) {
  extern __shared__ uint32_t cnt[];
  for (int p = threadIdx.x; p < nparts; p += blockDim.x) cnt[p] = 0;
  __syncthreads();
  const int64_t lo = (int64_t)blockIdx.x * rows_per_block, hi = min(n, lo + rows_per_block);
  for (int64_t r0 = lo + threadIdx.x; r0 < hi; r0 += (int64_t)PART_U * blockDim.x) {
    uint64_t w[PART_U][MAX_KEYS + 1];
    int64_t row[PART_U];
    bool ok[PART_U];
    load_keys_u(keys, key_null, r0, lo, hi, T.nkw, w, row, ok);
#pragma unroll
    for (int u = 0; u < PART_U; u++)
      if (ok[u]) atomicAdd(&cnt[(key_hash(w[u], T.nkw) & T.mask) >> T.part_shift], 1u);
  }
  __syncthreads();
  for (int p = threadIdx.x; p < nparts; p += blockDim.x) hist[(uint64_t)p * gridDim.x + blockIdx.x] = cnt[p];
}

__global__ void __launch_bounds__(256) join_part_scatter_kernel(DevJoinTable T, const uint64_t* keys,
                                                                const uint8_t* key_null, int64_t n,
                                                                int64_t rows_per_block, int nparts,
                                                                const uint32_t* offsets, uint32_t* srow,
                                                                uint64_t* skeys) {
  extern __shared__ uint32_t cur[];
  for (int p = threadIdx.x; p < nparts; p += blockDim.x) cur[p] = offsets[(uint64_t)p * gridDim.x + blockIdx.x];
  __syncthreads();
  const int64_t lo = (int64_t)blockIdx.x * rows_per_block, hi = min(n, lo + rows_per_block);
  const int nkw = T.nkw;
  for (int64_t r0 = lo + threadIdx.x; r0 < hi; r0 += (int64_t)PART_U * blockDim.x) {
    uint64_t w[PART_U][MAX_KEYS + 1];
    int64_t row[PART_U];
    bool ok[PART_U];
    load_keys_u(keys, key_null, r0, lo, hi, nkw, w, row, ok);
#pragma unroll
    for (int u = 0; u < PART_U; u++) {
      if (!ok[u]) continue;
      const uint64_t h = key_hash(w[u], nkw);
      const uint32_t at = atomicAdd(&cur[(h & T.mask) >> T.part_shift], 1u);
      srow[at] = (uint32_t)row[u];
      for (int k = 0; k < nkw; k++) skeys[(uint64_t)at * nkw + k] = w[u][k];
    }
  }
}

// One workgroup per partition: insert its keys into an LDS image of the
// partition's slots and Bloom words, check key uniqueness, then write both
// out with coalesced stores. flags bit 0: a duplicate key; bit 1: a partition
// with too many keys to keep an empty slot (the caller rebuilds unpartitioned).
__global__ void __launch_bounds__(256) join_part_build_kernel(DevJoinTable T, const uint32_t* offsets, int nblocks,
                                                              const uint32_t* srow,
                                                              const uint64_t* skeys, uint64_t* bloom,
                                                              uint32_t* flags) {
  extern __shared__ uint64_t lds[];
  const int p = blockIdx.x;
  const int nkw = T.nkw, stride = 1 + nkw;
  const uint32_t S = (uint32_t)T.wrap + 1u;
  const uint32_t W = bloom ? (uint32_t)T.bloom_mask + 1u : 0u;
  uint64_t* sl = lds;
  uint64_t* bw = lds + (uint64_t)S * stride;
  const uint32_t lo = offsets[(uint64_t)p * nblocks];
  const uint32_t hi = offsets[(uint64_t)(p + 1) * nblocks];  // the scan's last entry is the total
  if (hi - lo >= S) {  // uniform: no empty slot would be left
    if (threadIdx.x == 0) atomicOr(flags, 2u);
    return;
  }
  for (uint32_t i = threadIdx.x; i < S * stride; i += blockDim.x) sl[i] = 0;
  for (uint32_t i = threadIdx.x; i < W; i += blockDim.x) bw[i] = 0;
  __syncthreads();
  for (uint32_t e = lo + threadIdx.x; e < hi; e += blockDim.x) {
    uint64_t w[MAX_KEYS + 1];
    load_key(skeys, e, nkw, w);
    const uint64_t h = key_hash(w, nkw);
    const uint64_t hdr = ((h >> 32) << 32) | (uint64_t)(srow[e] + 1u);
    uint32_t pos = (uint32_t)(h & T.wrap);
    for (;;) {  // ends: fewer keys than slots
      if (atomicCAS((unsigned long long*)&sl[pos * stride], 0ull, (unsigned long long)hdr) == 0ull) {
        for (int k = 0; k < nkw; k++) sl[pos * stride + 1 + k] = w[k];
        break;
      }
      pos = (pos + 1) & (S - 1);
    }
    if (W) {
      const uint64_t g = bloom_mix(h);
      atomicOr((unsigned long long*)&bw[(g >> 18) & T.bloom_mask], (unsigned long long)bloom_bits(g));
    }
  }
  __syncthreads();
  // a key is a duplicate if an equal key sits earlier on its probe path
  bool dup = false;
  for (uint32_t e = lo + threadIdx.x; e < hi && !dup; e += blockDim.x) {
    uint64_t w[MAX_KEYS + 1];
    load_key(skeys, e, nkw, w);
    const uint64_t h = key_hash(w, nkw);
    const uint64_t me = ((h >> 32) << 32) | (uint64_t)(srow[e] + 1u);
    uint32_t pos = (uint32_t)(h & T.wrap);
    for (;;) {
      const uint64_t hd = sl[pos * stride];
      if (hd == me) break;
      if ((hd >> 32) == (h >> 32)) {
        bool eq = true;
        for (int k = 0; k < nkw; k++) eq = eq && sl[pos * stride + 1 + k] == w[k];
        if (eq) {
          dup = true;
          break;
        }
      }
      pos = (pos + 1) & (S - 1);
    }
  }
  if (dup) atomicOr(flags, 1u);
  uint64_t* out = T.slots + (uint64_t)p * S * stride;
  for (uint32_t i = threadIdx.x; i < S * stride; i += blockDim.x) out[i] = sl[i];
  if (W)
    for (uint32_t i = threadIdx.x; i < W; i += blockDim.x) bloom[(uint64_t)p * W + i] = bw[i];
}

// --------------------------------------------------------------- agg tables
__global__ void agg_rehash_kernel(const uint64_t* old_slots, uint64_t old_cap, int nkw, uint64_t* new_slots,
                                  uint64_t new_mask, AccMove mv) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < (int64_t)old_cap; p += stride) {
    const uint64_t* s = old_slots + (uint64_t)p * (1 + nkw);
    const uint64_t hdr = s[0];
    if (hdr == 0) continue;
    uint64_t w[MAX_KEYS + 1];
#pragma unroll
    for (int k = 0; k < MAX_KEYS + 1; k++) w[k] = k < nkw ? s[1 + k] : 0;
    const uint64_t h = key_hash(w, nkw);
    uint64_t pos = h & new_mask;
    for (;;) {
      uint64_t* t = new_slots + pos * (1 + nkw);
      if (atomicCAS((unsigned long long*)t, 0ull, (unsigned long long)hdr) == 0ull) {
#pragma unroll
        for (int k = 0; k < MAX_KEYS + 1; k++)
          if (k < nkw) t[1 + k] = w[k];
        for (int a = 0; a < mv.n; a++) ((uint64_t*)mv.dst[a])[pos] = ((const uint64_t*)mv.src[a])[p];
        break;
      }
      pos = (pos + 1) & new_mask;
    }
  }
}

__global__ void slot_flags_kernel(const uint64_t* slots, uint64_t cap, int nkw, uint8_t* flags) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < (int64_t)cap; p += stride)
    flags[p] = slots[(uint64_t)p * (1 + nkw)] != 0;
}

// Decode group keys and finish aggregates for output rows j (slot = idx[j]).
__global__ void agg_emit_kernel(const uint64_t* slots, int nkw, const int64_t* idx, int64_t n, EmitPlan E) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride) {
    const int64_t slot = idx[j];
    const uint64_t* s = slots + (uint64_t)slot * (1 + nkw);
    const uint64_t nullmask = E.key_nullable ? s[1 + E.nkeys] : 0;
    for (int k = 0; k < E.nkeys; k++) {
      uint64_t w = s[1 + k];
      if (((E.key_f64 >> k) & 1) && f64_bits_nan(w)) w = CANONICAL_NAN;
      if (E.key_width[k] == 8) ((uint64_t*)E.key_out[k])[j] = w;
      else if (E.key_width[k] == 4) ((uint32_t*)E.key_out[k])[j] = (uint32_t)w;
      else ((uint8_t*)E.key_out[k])[j] = (uint8_t)w;
      if (E.key_valid[k]) E.key_valid[k][j] = ((nullmask >> k) & 1) == 0;
    }
    for (int a = 0; a < E.nout; a++) {
      const uint64_t v = ((const uint64_t*)E.state[a])[slot];
      const int64_t cnt = E.count_state[a] ? ((const int64_t*)E.count_state[a])[slot] : 1;
      uint64_t out = v;
      switch (E.finish[a]) {
        case FIN_AVG: out = f2u(cnt == 0 ? 0.0 : u2f(v) / (double)cnt); break;
        case FIN_I32: out = (uint64_t)(uint32_t)v; break;
        case FIN_RANK: out = (cnt > 0 && v < E.rank_n[a]) ? (uint64_t)E.rank_code[a][v] : 0; break;
        default: break;
      }
      if (E.out_width[a] == 8) ((uint64_t*)E.out[a])[j] = out;
      else ((uint32_t*)E.out[a])[j] = (uint32_t)out;
      if (E.out_valid[a]) E.out_valid[a][j] = cnt > 0;
    }
  }
}

// ------------------------------------------------------- ordered compaction
constexpr int CT = 256;          // threads
constexpr int CE = 4;            // elements per thread
constexpr int CB = CT * CE;      // elements per block

__global__ void flag_counts_kernel(const uint8_t* flags, int64_t n, uint32_t* block_counts) {
  __shared__ uint32_t red[CT / 64];
  const int64_t base = (int64_t)blockIdx.x * CB + (int64_t)threadIdx.x * CE;
  uint32_t c = 0;
#pragma unroll
  for (int e = 0; e < CE; e++) c += (base + e < n) ? (flags[base + e] != 0) : 0;
  for (int o = 32; o > 0; o >>= 1) c += __shfl_down(c, o);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int w = 0; w < CT / 64; w++) t += red[w];
    block_counts[blockIdx.x] = t;
  }
}

// exclusive scan of block counts (single workgroup, chunked); total at offsets[nb]
// Exclusive scan of nb u32 counts into u64 offsets, offsets[nb] = total; one
// 1024-thread workgroup, each thread summing SCAN_PER consecutive counts so a
// round of the workgroup scan covers 16K counts.
constexpr int SCAN_PER = 16;
__global__ void scan_counts_kernel(const uint32_t* counts, int64_t nb, uint64_t* offsets) {
  __shared__ uint64_t buf[1024];
  uint64_t carry = 0;
  for (int64_t base = 0; base < nb; base += 1024 * SCAN_PER) {
    const int64_t first = base + (int64_t)threadIdx.x * SCAN_PER;
    uint64_t local = 0;
    for (int k = 0; k < SCAN_PER; k++)
      if (first + k < nb) local += counts[first + k];
    buf[threadIdx.x] = local;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
      const uint64_t t = threadIdx.x >= (unsigned)o ? buf[threadIdx.x - o] : 0;
      __syncthreads();
      buf[threadIdx.x] += t;
      __syncthreads();
    }
    uint64_t run = carry + buf[threadIdx.x] - local;
    for (int k = 0; k < SCAN_PER; k++)
      if (first + k < nb) {
        offsets[first + k] = run;
        run += counts[first + k];
      }
    const uint64_t total = buf[1023];
    __syncthreads();
    carry += total;
  }
  if (threadIdx.x == 0) offsets[nb] = carry;
}

__global__ void flag_scatter_kernel(const uint8_t* flags, int64_t n, const uint64_t* offsets, const uint32_t* sel,
                                    void* out, int out_u32) {
  __shared__ uint32_t warp_tot[CT / 64];
  const int64_t base = (int64_t)blockIdx.x * CB + (int64_t)threadIdx.x * CE;
  uint32_t f[CE];
  uint32_t c = 0;
#pragma unroll
  for (int e = 0; e < CE; e++) {
    f[e] = (base + e < n) ? (flags[base + e] != 0) : 0;
    c += f[e];
  }
  // inclusive scan within the wave
  const int lane = threadIdx.x & 63;
  uint32_t incl = c;
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t t = __shfl_up(incl, o);
    if (lane >= o) incl += t;
  }
  if (lane == 63) warp_tot[threadIdx.x >> 6] = incl;
  __syncthreads();
  uint32_t wbase = 0;
  for (int w = 0; w < (int)(threadIdx.x >> 6); w++) wbase += warp_tot[w];
  uint64_t pos = offsets[blockIdx.x] + wbase + incl - c;
#pragma unroll
  for (int e = 0; e < CE; e++) {
    if (f[e]) {
      const int64_t i = base + e;
      const uint64_t v = sel ? (uint64_t)sel[i] : (uint64_t)i;
      if (out_u32) ((uint32_t*)out)[pos] = (uint32_t)v;
      else ((int64_t*)out)[pos] = (int64_t)v;
      pos++;
    }
  }
}

// --------------------------------------------- partitioned hash aggregation
// Key words of materialized record i (NULL keys: word 0 plus the null-mask word).
__device__ __forceinline__ void record_keys(const AggRecords& R, int64_t i, uint64_t (&w)[MAX_KEYS + 1]) {
  uint64_t nullmask = 0;
#pragma unroll
  for (int k = 0; k < MAX_KEYS + 1; k++) w[k] = 0;
#pragma unroll
  for (int k = 0; k < MAX_KEYS; k++) {
    if (k >= R.nkeys) break;
    const bool ok = !R.key_ok[k] || R.key_ok[k][i];
    w[k] = ok ? R.key_val[k][i] : 0;
    if (!ok) nullmask |= 1ull << k;
  }
  if (R.key_nullable) set_word(w, R.nkeys, nullmask);
}

// Both passes run PART_PASS_THREADS threads per workgroup: a workgroup owns
// one slice of the per-partition offsets, so its threads (not more
// workgroups, which would multiply the histogram) carry the records in flight.
constexpr int PART_PASS_THREADS = 1024;

__global__ void __launch_bounds__(PART_PASS_THREADS) agg_part_hist_kernel(AggRecords R, int64_t n, uint64_t mask, int part_shift,
                                                            int64_t rows_per_block, int nparts, uint32_t* hist) {
  extern __shared__ uint32_t cnt[];
  for (int p = threadIdx.x; p < nparts; p += blockDim.x) cnt[p] = 0;
  __syncthreads();
  const int64_t lo = (int64_t)blockIdx.x * rows_per_block, hi = min(n, lo + rows_per_block);
  for (int64_t i = lo + threadIdx.x; i < hi; i += blockDim.x) {
    uint64_t w[MAX_KEYS + 1];
    record_keys(R, i, w);
    atomicAdd(&cnt[(key_hash(w, R.nkw) & mask) >> part_shift], 1u);
  }
  __syncthreads();
  for (int p = threadIdx.x; p < nparts; p += blockDim.x) hist[(uint64_t)p * gridDim.x + blockIdx.x] = cnt[p];
}

__global__ void __launch_bounds__(PART_PASS_THREADS) agg_part_scatter_kernel(AggRecords R, int64_t n, uint64_t mask, int part_shift,
                                                               int64_t rows_per_block, int nparts,
                                                               const uint32_t* offsets, uint64_t* rkeys,
                                                               uint64_t* rargs, uint32_t* rvalid) {
  extern __shared__ uint32_t cur[];
  for (int p = threadIdx.x; p < nparts; p += blockDim.x) cur[p] = offsets[(uint64_t)p * gridDim.x + blockIdx.x];
  __syncthreads();
  const int64_t lo = (int64_t)blockIdx.x * rows_per_block, hi = min(n, lo + rows_per_block);
  for (int64_t i = lo + threadIdx.x; i < hi; i += blockDim.x) {
    uint64_t w[MAX_KEYS + 1];
    record_keys(R, i, w);
    const uint32_t at = atomicAdd(&cur[(key_hash(w, R.nkw) & mask) >> part_shift], 1u);
    for (int k = 0; k < R.nkw; k++) rkeys[(uint64_t)k * n + at] = w[k];
    uint32_t vm = 0;
    for (int a = 0; a < R.nargs; a++) {
      rargs[(uint64_t)a * n + at] = R.arg_val[a][i];
      if (!R.arg_ok[a] || R.arg_ok[a][i]) vm |= 1u << a;
    }
    if (rvalid) rvalid[at] = vm;
  }
}

// Merge an accumulator state into an LDS state word.
__device__ __forceinline__ void acc_merge_lds(int kind, uint64_t* w, uint64_t v) {
  if (kind == ACC_COUNT) atomicAdd((unsigned long long*)w, (unsigned long long)v);
  else acc_apply_lds(kind, w, v);
}

// Find or insert the group of key words w (table hash h) in an LDS table of S
// slots (headers hdr, keys kw). Returns the slot, or -1 when the table is
// full. A new group also counts itself in *groups; past `limit` groups it is
// still completed (lanes may wait on its header) but -1 is returned: a table
// stops at its fill limit, before linear probes over a nearly full table get
// long.
__device__ __forceinline__ int lds_group(uint64_t* hdr, uint64_t* kw, int S, int nkw,
                                         const uint64_t (&w)[MAX_KEYS + 1], uint64_t h, uint32_t* groups,
                                         uint32_t limit) {
  const uint64_t tag = (h & ~3ull) | 2ull;
  int pos = (int)(h & (uint64_t)(S - 1));
  int probes = 0;
  while (probes < S) {
    uint64_t c = __hip_atomic_load(&hdr[pos], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (c == 0) {
      const uint64_t prev = atomicCAS((unsigned long long*)&hdr[pos], 0ull, 1ull);
      if (prev == 0) {
        for (int k = 0; k < nkw; k++) kw[(uint64_t)pos * nkw + k] = w[k];
        __hip_atomic_store(&hdr[pos], tag, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
        return atomicAdd(groups, 1u) < limit ? pos : -1;
      }
      c = prev;
    }
    if (c == 1) continue;  // being written by another lane: read again
    if (c == tag) {
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
      bool eq = true;
      for (int k = 0; k < nkw; k++) eq = eq && kw[(uint64_t)pos * nkw + k] == w[k];
      if (eq) return pos;
    }
    pos = (pos + 1) & (S - 1);
    probes++;
  }
  return -1;
}


// LDS table layout: hdr[S], kw[S * nkw], acc[nacc * S] (u64 each)
size_t agg_part_lds(const PartAggPlan& A) { return (size_t)A.S * (1 + A.nkw + A.nacc) * 8; }

// Write an LDS table to a table region (slots [base, base + S) of `slots` and
// accumulator arrays `acc_data`); count its groups into flags.
__device__ __forceinline__ void part_write(const PartAggPlan& A, const uint64_t* hdr, const uint64_t* kw,
                                           const uint64_t* acc, uint64_t* slots, void* const* acc_data, uint64_t base,
                                           bool count) {
  const int S = A.S, stride = 1 + A.nkw;
  uint32_t groups = 0;
  for (int i = threadIdx.x; i < S; i += blockDim.x) {
    const uint64_t h = hdr[i];
    groups += h != 0;
    slots[(base + i) * stride] = h;
    for (int k = 0; k < A.nkw; k++) slots[(base + i) * stride + 1 + k] = kw[(uint64_t)i * A.nkw + k];
    for (int s2 = 0; s2 < A.nacc; s2++) ((uint64_t*)acc_data[s2])[base + i] = acc[(uint64_t)s2 * S + i];
  }
  if (count) {
    for (int o = 32; o > 0; o >>= 1) groups += __shfl_down(groups, o);
    if (__lane_id() == 0 && groups) atomicAdd((unsigned long long*)A.flags, (unsigned long long)groups);
  }
}

__device__ __forceinline__ void part_init(const PartAggPlan& A, uint64_t* hdr, uint64_t* acc) {
  for (int i = threadIdx.x; i < A.S; i += blockDim.x) hdr[i] = 0;
  for (int i = threadIdx.x; i < A.S * A.nacc; i += blockDim.x) acc[i] = acc_init_bits(A.acc[i / A.S].kind);
  __syncthreads();
}

// One workgroup per chunk of a partition's records.
__global__ void __launch_bounds__(256) agg_part_chunk_kernel(PartAggPlan A, const PartChunk* chunks,
                                                             const uint64_t* rkeys, const uint64_t* rargs,
                                                             const uint32_t* rvalid, int64_t n, uint64_t* partial) {
  extern __shared__ uint64_t lds[];
  __shared__ uint32_t table_full;  // the batch is aggregated again: stop early
  __shared__ uint32_t fail_at;     // the first record that found the table full
  __shared__ uint32_t groups;      // groups in the table
  const PartChunk C = chunks[blockIdx.x];
  const int S = A.S;
  uint64_t* hdr = lds;
  uint64_t* kw = lds + S;
  uint64_t* acc = kw + (uint64_t)S * A.nkw;
  if (threadIdx.x == 0) {
    table_full = 0;
    fail_at = C.hi;
    groups = 0;
  }
  part_init(A, hdr, acc);
  bool full = false;
  for (uint32_t e = C.lo + threadIdx.x; e < C.hi; e += blockDim.x) {
    if (__hip_atomic_load(&table_full, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) break;
    uint64_t w[MAX_KEYS + 1];
#pragma unroll
    for (int k = 0; k < MAX_KEYS + 1; k++) w[k] = k < A.nkw ? rkeys[(uint64_t)k * n + e] : 0;
    const int pos = lds_group(hdr, kw, S, A.nkw, w, key_hash(w, A.nkw), &groups, part_fill_limit((uint32_t)S));
    if (pos < 0) {
      // the table reached its fill limit: the overflow reruns the batch, so
      // nothing more of this chunk is needed
      full = true;
      __hip_atomic_store(&table_full, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      atomicMin(&fail_at, e);
      break;
    }
    const uint32_t vm = rvalid ? rvalid[e] : 0xFFFFFFFFu;
    for (int a = 0; a < A.nargs; a++) {
      if (!((vm >> a) & 1u)) continue;  // NULL arguments are skipped by every accumulator
      const uint64_t v = rargs[(uint64_t)a * n + e];
      for (int s2 = 0; s2 < A.nacc; s2++)
        if (A.acc[s2].arg == a) acc_apply_lds(A.acc[s2].kind, acc + (uint64_t)s2 * S + pos, v);
    }
    for (int s2 = 0; s2 < A.nacc; s2++)
      if (A.acc[s2].arg < 0) atomicAdd((unsigned long long*)(acc + (uint64_t)s2 * S + pos), 1ull);
  }
  if (__any(full) && __lane_id() == 0) atomicOr(A.flags + 2, 2u);
  __syncthreads();
  // partitions whose table filled up (a split partition is counted by its
  // merge), each with its projected groups: the fill limit's groups after the
  // records before fail_at, scaled to all of the partition's records
  if (threadIdx.x == 0 && table_full && C.partial == 0xFFFFFFFFu) {
    atomicAdd(A.flags + 3, 1u);
    const uint64_t seen = max(1u, fail_at - C.lo);
    atomicAdd((unsigned long long*)(A.flags + 4),
              (unsigned long long)((uint64_t)part_fill_limit((uint32_t)S) * (C.hi - C.lo) / seen));
  }
  if (C.partial == 0xFFFFFFFFu) {
    part_write(A, hdr, kw, acc, A.slots, A.acc_data, (uint64_t)C.part * S, true);
  } else {  // a partial table: [S] headers, [S * nkw] keys, [nacc * S] states
    uint64_t* q = partial + (uint64_t)C.partial * S * (1 + A.nkw + A.nacc);
    for (int i = threadIdx.x; i < S * (1 + A.nkw + A.nacc); i += blockDim.x) q[i] = lds[i];
  }
}

// One workgroup per partition that was split into several chunks: merge the
// chunks' partial tables first[j] .. first[j+1]-1 and write the partition.
// The (partial, slot) pairs are spread over all threads with no barrier
// between partials (inserts and merges into the LDS table are atomic), and
// each thread requests the headers and keys of MERGE_U slots before using
// any: a partition of a hot key can have ~100 partials, and reading them one
// dependent load at a time left one workgroup running long after the rest.
constexpr int MERGE_THREADS = 1024;
constexpr int MERGE_U = 4;
__global__ void __launch_bounds__(MERGE_THREADS) agg_part_merge_kernel(PartAggPlan A, const uint32_t* parts,
                                                                       const uint32_t* first, const uint64_t* partial) {
  extern __shared__ uint64_t lds[];
  __shared__ uint32_t table_full;  // the batch is aggregated again: stop early
  __shared__ uint32_t groups;      // groups in the table
  const int S = A.S, nkw = A.nkw;
  uint64_t* hdr = lds;
  uint64_t* kw = lds + S;
  uint64_t* acc = kw + (uint64_t)S * nkw;
  if (threadIdx.x == 0) {
    table_full = 0;
    groups = 0;
  }
  part_init(A, hdr, acc);
  bool full = false;
  const size_t words = (size_t)S * (1 + nkw + A.nacc);
  const uint64_t q0 = first[blockIdx.x];
  const uint64_t total = (uint64_t)(first[blockIdx.x + 1] - q0) * (uint64_t)S;
  const int log_s = __builtin_ctz((unsigned)S);
  for (uint64_t base = threadIdx.x; base < total; base += (uint64_t)MERGE_U * blockDim.x) {
    if (__hip_atomic_load(&table_full, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) break;
    const uint64_t* t[MERGE_U];
    int slot[MERGE_U];
    uint64_t h[MERGE_U];
    uint64_t w[MERGE_U][MAX_KEYS + 1];
#pragma unroll
    for (int u = 0; u < MERGE_U; u++) {
      const uint64_t e = base + (uint64_t)u * blockDim.x;
      const bool in = e < total;
      const uint64_t ee = in ? e : base;
      t[u] = partial + (q0 + (ee >> log_s)) * words;
      slot[u] = (int)(ee & (uint64_t)(S - 1));
      h[u] = in ? t[u][slot[u]] : 0;
#pragma unroll
      for (int k = 0; k < MAX_KEYS + 1; k++) w[u][k] = k < nkw ? t[u][S + (uint64_t)slot[u] * nkw + k] : 0;
    }
#pragma unroll
    for (int u = 0; u < MERGE_U; u++) {
      if (h[u] == 0 || full) continue;
      const int pos = lds_group(hdr, kw, S, nkw, w[u], key_hash(w[u], nkw), &groups, part_fill_limit((uint32_t)S));
      if (pos < 0) {
        full = true;
        __hip_atomic_store(&table_full, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        continue;
      }
      const uint64_t* ta = t[u] + S + (uint64_t)S * nkw;
      for (int s2 = 0; s2 < A.nacc; s2++) {
        const uint64_t v = ta[(uint64_t)s2 * S + slot[u]];
        if (v != acc_init_bits(A.acc[s2].kind)) acc_merge_lds(A.acc[s2].kind, acc + (uint64_t)s2 * S + pos, v);
      }
    }
  }
  if (__any(full) && __lane_id() == 0) atomicOr(A.flags + 2, 2u);
  __syncthreads();
  if (threadIdx.x == 0 && table_full) {  // no projection here: at least twice the table
    atomicAdd(A.flags + 3, 1u);
    atomicAdd((unsigned long long*)(A.flags + 4), (unsigned long long)(2 * S));
  }
  part_write(A, hdr, kw, acc, A.slots, A.acc_data, (uint64_t)parts[blockIdx.x] * S, true);
}

// ---------------------------------------------- exclusive scan of u32 counts
// Three passes: per-workgroup sums of SCAN_B elements, one workgroup scanning
// those sums, then each workgroup scanning its elements from its base.
constexpr int SCAN_T = 256;
constexpr int SCAN_E = 16;
constexpr int SCAN_B = SCAN_T * SCAN_E;

__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* red, uint32_t* total) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  uint32_t incl = v;
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t t = __shfl_up(incl, o);
    if (lane >= o) incl += t;
  }
  if (lane == 63) red[wv] = incl;
  __syncthreads();
  uint32_t base = 0, all = 0;
  for (int w = 0; w < SCAN_T / 64; w++) {
    if (w < wv) base += red[w];
    all += red[w];
  }
  *total = all;
  __syncthreads();
  return base + incl - v;
}

// Scan input i: a u32 count, or (popc) the set bits of u64 word i.
__device__ __forceinline__ uint32_t scan_in(const void* in, int64_t i, int popc) {
  return popc ? (uint32_t)__popcll(((const uint64_t*)in)[i]) : ((const uint32_t*)in)[i];
}

__global__ void __launch_bounds__(SCAN_T) scan_u32_reduce_kernel(const void* in, int popc, int64_t n, uint32_t* sums) {
  __shared__ uint32_t red[SCAN_T / 64];
  const int64_t base = (int64_t)blockIdx.x * SCAN_B + (int64_t)threadIdx.x * SCAN_E;
  uint32_t c = 0;
#pragma unroll
  for (int e = 0; e < SCAN_E; e++) c += (base + e < n) ? scan_in(in, base + e, popc) : 0u;
  uint32_t total;
  block_excl_scan(c, red, &total);
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// in place: sums[i] <- exclusive prefix, sums[nb] <- total
__global__ void __launch_bounds__(SCAN_T) scan_u32_top_kernel(uint32_t* sums, int64_t nb) {
  __shared__ uint32_t red[SCAN_T / 64];
  uint32_t carry = 0;
  for (int64_t b0 = 0; b0 < nb; b0 += SCAN_T) {
    const int64_t i = b0 + threadIdx.x;
    const uint32_t v = i < nb ? sums[i] : 0u;
    uint32_t total;
    const uint32_t ex = block_excl_scan(v, red, &total);
    if (i < nb) sums[i] = carry + ex;
    carry += total;
  }
  if (threadIdx.x == 0) sums[nb] = carry;
}

__global__ void __launch_bounds__(SCAN_T) scan_u32_down_kernel(const void* in, int popc, int64_t n,
                                                               const uint32_t* sums, uint32_t* out) {
  __shared__ uint32_t red[SCAN_T / 64];
  const int64_t base = (int64_t)blockIdx.x * SCAN_B + (int64_t)threadIdx.x * SCAN_E;
  uint32_t v[SCAN_E];
  uint32_t c = 0;
#pragma unroll
  for (int e = 0; e < SCAN_E; e++) {
    v[e] = (base + e < n) ? scan_in(in, base + e, popc) : 0u;
    c += v[e];
  }
  uint32_t total;
  uint32_t run = sums[blockIdx.x] + block_excl_scan(c, red, &total);
#pragma unroll
  for (int e = 0; e < SCAN_E; e++) {
    if (base + e < n) out[base + e] = run;
    run += v[e];
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) out[n] = sums[gridDim.x];
}

// ------------------------------------------------------------------- TopN
// Sort words: lexicographic u64 words, then row index ascending (lex_less).

__global__ void sort_keys_kernel(Plan P, SortKeys K, int64_t n, uint64_t* words) {
  const Code ins{(CInstr*)P.ins, P.nins - 1};
  const int nw = K.nwords;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    uint32_t rows[MAX_STAGES];
#pragma unroll
    for (int s = 0; s < MAX_STAGES; s++) rows[s] = 0;
    rows[0] = i;
    uint64_t* out = words + (uint64_t)i * (nw + 1);
    int wi = 0;
    for (int k = 0; k < K.nkeys; k++) {
      uint64_t v;
      const bool ok = eval(ins, K.prog[k], rows, v);
      uint64_t word;
      if (K.kind[k] == SK_F64) {  // f64::total_cmp order
        const int64_t b = (int64_t)v;
        word = (uint64_t)(b ^ (int64_t)(((uint64_t)(b >> 63)) >> 1)) ^ 0x8000000000000000ull;
      } else if (K.kind[k] == SK_INT) {
        word = v ^ 0x8000000000000000ull;
      } else {
        word = v;
      }
      if (!K.asc[k]) word = ~word;
      if (K.nullable[k]) {
        const uint64_t nflag = ok ? (K.nulls_first[k] ? 1 : 0) : (K.nulls_first[k] ? 0 : 1);
        out[wi++] = nflag;
        out[wi++] = ok ? word : 0;
      } else {
        out[wi++] = word;
      }
    }
    out[nw] = (uint64_t)i;
  }
}

// Bitonic-sort tiles of TILE entries (nw+1 words each) in LDS; keep the first
// `keep` entries of every tile. A short tile sorts only the power of two that
// covers its entries (the +inf padding beyond them would stay in place).
// Lexicographic order of two W-word entries (key words, then the row index).
template <int W>
__device__ __forceinline__ bool lex_less(const uint64_t (&a)[W], const uint64_t (&b)[W]) {
  bool lt = false, eq = true;
#pragma unroll
  for (int k = 0; k < W; k++) {
    lt = lt || (eq && a[k] < b[k]);
    eq = eq && a[k] == b[k];
  }
  return lt;
}

template <int W>
__global__ void __launch_bounds__(256) topn_tile_kernel(const uint64_t* in, int64_t n, int keep, uint64_t* out) {
  extern __shared__ uint64_t sh[];
  const int64_t tile0 = (int64_t)blockIdx.x * TOPN_TILE;
  const int cnt = (int)((n - tile0) < TOPN_TILE ? (n - tile0) : TOPN_TILE);
  // the tile's entries are contiguous words: load them flat (coalesced), pad with +inf
  for (int e = threadIdx.x; e < TOPN_TILE * W; e += blockDim.x)
    sh[e] = e < cnt * W ? in[(uint64_t)tile0 * W + e] : ~0ull;
  __syncthreads();
  int span = 2;
  while (span < cnt) span <<= 1;
  for (int size = 2; size <= span; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      const int lg = __builtin_ctz(stride);
      for (int t = threadIdx.x; t < span / 2; t += blockDim.x) {
        const int lo = ((t >> lg) << (lg + 1)) + (t & (stride - 1));
        const int hi = lo + stride;
        const bool up = ((lo & size) == 0);
        uint64_t a[W], b[W];
#pragma unroll
        for (int k = 0; k < W; k++) {
          a[k] = sh[lo * W + k];
          b[k] = sh[hi * W + k];
        }
        const bool swap = up ? lex_less(b, a) : lex_less(a, b);
        if (swap) {
#pragma unroll
          for (int k = 0; k < W; k++) {
            sh[lo * W + k] = b[k];
            sh[hi * W + k] = a[k];
          }
        }
      }
      __syncthreads();
    }
  }
  // the first `keep` entries, flat (a short tile's missing entries are the +inf padding)
  const int m = keep < cnt ? keep : cnt;
  uint64_t* dst = out + (uint64_t)blockIdx.x * keep * W;
  for (int e = threadIdx.x; e < keep * W; e += blockDim.x) dst[e] = e < m * W ? sh[e] : ~0ull;
}

// Rank selection for a small limit: tiles of TOPN_RANK_TILE entries; the
// rank of an entry is the number of entries of its tile that order before it
// (lex_less, ties by position, so every entry gets its own rank), counted by
// TOPN_RANK_SPLIT threads over disjoint quarters of the tile and summed; the
// entry is written at its rank when that is below `keep`. A short tile's
// missing ranks are the +inf padding. No sorting network: two barriers, and
// the tile's entries are read as LDS broadcasts. W (words per entry) is a
// template parameter, so the comparison has no per-word branches.
template <int W>
__global__ void __launch_bounds__(TOPN_RANK_TILE * TOPN_RANK_SPLIT) topn_rank_kernel(const uint64_t* in, int64_t n,
                                                                                      int keep, uint64_t* out) {
  __shared__ uint64_t sh[TOPN_RANK_TILE * W];
  __shared__ int part[TOPN_RANK_SPLIT][TOPN_RANK_TILE];
  const int64_t tile0 = (int64_t)blockIdx.x * TOPN_RANK_TILE;
  const int cnt = (int)((n - tile0) < TOPN_RANK_TILE ? (n - tile0) : TOPN_RANK_TILE);
  for (int e = threadIdx.x; e < cnt * W; e += blockDim.x) sh[e] = in[(uint64_t)tile0 * W + e];
  __syncthreads();
  const int i = (int)threadIdx.x % TOPN_RANK_TILE;  // the entry
  const int q = (int)threadIdx.x / TOPN_RANK_TILE;  // its quarter of the tile
  if (i < cnt) {
    uint64_t mine[W];
#pragma unroll
    for (int k = 0; k < W; k++) mine[k] = sh[i * W + k];
    const int j0 = cnt * q / TOPN_RANK_SPLIT, j1 = cnt * (q + 1) / TOPN_RANK_SPLIT;
    int rank = 0;
#pragma unroll 4
    for (int j = j0; j < j1; j++) {
      bool lt = false, eq = true;
#pragma unroll
      for (int k = 0; k < W; k++) {
        const uint64_t x = sh[j * W + k];
        lt = lt || (eq && x < mine[k]);
        eq = eq && x == mine[k];
      }
      rank += (lt || (eq && j < i)) ? 1 : 0;
    }
    part[q][i] = rank;
  }
  __syncthreads();
  uint64_t* dst = out + (uint64_t)blockIdx.x * keep * W;
  if (q == 0 && i < cnt) {
    int rank = 0;
#pragma unroll
    for (int p = 0; p < TOPN_RANK_SPLIT; p++) rank += part[p][i];
    if (rank < keep) {
#pragma unroll
      for (int k = 0; k < W; k++) dst[(uint64_t)rank * W + k] = sh[i * W + k];
    }
  }
  for (int e = cnt * W + (int)threadIdx.x; e < keep * W; e += blockDim.x) dst[e] = ~0ull;
}

// ---------------------------------------------------------- full sort
// Merge sort of entries of W words (lexicographic over all W words; the
// last word is the row index, so entries differ): LDS-sorted tiles of
// TOPN_TILE entries, then passes merging pairs of sorted runs of length L
// (merge path: each workgroup finds where its 1024 outputs split between
// the two runs, stages both pieces in LDS and merges them there).
constexpr int MERGE_OUT = 1024;  // outputs per workgroup in a merge pass

template <int W>
__device__ __forceinline__ bool words_le(const uint64_t* a, const uint64_t* b) {
  uint64_t x[W], y[W];
#pragma unroll
  for (int k = 0; k < W; k++) {
    x[k] = a[k];
    y[k] = b[k];
  }
  return !lex_less(y, x);
}

// Entries of A placed among the first k outputs of merge(A, B); A goes first on ties.
template <int W>
__device__ __forceinline__ int64_t merge_corank(const uint64_t* A, int64_t na, const uint64_t* B, int64_t nb,
                                                int64_t k) {
  int64_t lo = k > nb ? k - nb : 0, hi = k < na ? k : na;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (words_le<W>(A + mid * W, B + (k - mid - 1) * W)) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// Where every workgroup's first output splits between its two runs: one
// thread per workgroup, all searches in parallel ahead of the merge.
template <int W>
__global__ void merge_split_kernel(const uint64_t* in, int64_t N, int64_t L, int64_t nblocks, int64_t* split) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nblocks) return;
  const int64_t o0 = b * MERGE_OUT;
  const int64_t ps = o0 / (2 * L) * (2 * L);
  const int64_t na = min(L, N - ps);
  const int64_t nb = max((int64_t)0, min(L, N - ps - L));
  const uint64_t* A = in + ps * W;
  split[b] = merge_corank<W>(A, na, A + na * W, nb, o0 - ps);
}

template <int W>
__global__ void __launch_bounds__(256) merge_pass_kernel(const uint64_t* in, uint64_t* out, int64_t N, int64_t L,
                                                         const int64_t* split) {
  extern __shared__ uint64_t sh[];
  const int64_t o0 = (int64_t)blockIdx.x * MERGE_OUT;
  const int64_t ps = o0 / (2 * L) * (2 * L);
  const int64_t na = min(L, N - ps);
  const int64_t nb = max((int64_t)0, min(L, N - ps - L));
  const uint64_t* A = in + ps * W;
  const uint64_t* B = A + na * W;
  const int64_t k0 = o0 - ps, k1 = min(k0 + MERGE_OUT, na + nb);
  // a block that ends its pair of runs ends with all of A; otherwise the
  // next block (same pair: MERGE_OUT divides 2L) starts where this one ends
  const int64_t i0 = split[blockIdx.x], i1 = k1 == na + nb ? na : split[blockIdx.x + 1];
  const int64_t j0 = k0 - i0, j1 = k1 - i1;
  const int ca = (int)(i1 - i0), cb = (int)(j1 - j0);
  uint64_t* sA = sh;
  uint64_t* sB = sh + (int64_t)ca * W;
  for (int e = threadIdx.x; e < ca * W; e += blockDim.x) sA[e] = A[i0 * W + e];
  for (int e = threadIdx.x; e < cb * W; e += blockDim.x) sB[e] = B[j0 * W + e];
  __syncthreads();
  // each thread merges PER outputs, noting which staged entry each one is;
  // then the workgroup writes the outputs out in order, coalesced
  __shared__ uint16_t pos[MERGE_OUT];
  constexpr int PER = MERGE_OUT / 256;
  const int cnt = ca + cb;
  const int t0 = (int)threadIdx.x * PER;
  if (t0 < cnt) {
    int ia = (int)merge_corank<W>(sA, ca, sB, cb, t0);
    int ib = t0 - ia;
    for (int e = 0; e < PER && t0 + e < cnt; e++) {
      const bool takeA = ib >= cb || (ia < ca && words_le<W>(sA + (int64_t)ia * W, sB + (int64_t)ib * W));
      pos[t0 + e] = (uint16_t)(takeA ? ia++ : ca + ib++);
    }
  }
  __syncthreads();
  uint64_t* dst = out + o0 * W;
  for (int e = threadIdx.x; e < cnt * W; e += blockDim.x) {
    const int r = e / W;
    dst[e] = sh[(int)pos[r] * W + (e - r * W)];
  }
}

// ------------------------------------------------------- validity formats
__global__ void bits_to_bytes_kernel(const uint64_t* bits, int64_t n, uint8_t* out) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    out[i] = (uint8_t)((bits[i >> 6] >> (i & 63)) & 1ull);
}

// one wave per 64-row word: the wave's ballot is the word
__global__ void bytes_to_bits_kernel(const uint8_t* valid, int64_t n, uint64_t* bits) {
  const int64_t nw = (n + 63) / 64;
  const int64_t waves = (int64_t)gridDim.x * (blockDim.x / 64);
  for (int64_t w = (int64_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64; w < nw; w += waves) {
    const int64_t r = w * 64 + (threadIdx.x & 63);
    const uint64_t m = __ballot(r < n && valid[r] != 0);
    if ((threadIdx.x & 63) == 0) bits[w] = m;
  }
}

// ------------------------------------------------------------ MLB1 bytes
// Row k of a batch is physical row sel[k] (or k without a selection).
__global__ void null_count_kernel(const uint8_t* valid, const uint32_t* sel, int64_t n, uint64_t* out) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  uint64_t c = 0;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride)
    c += valid[sel ? sel[k] : k] == 0;
  for (int o = 32; o > 0; o >>= 1) c += __shfl_down(c, o);
  if ((threadIdx.x & 63) == 0 && c) atomicAdd((unsigned long long*)out, (unsigned long long)c);
}

__global__ void mlb1_str_len_kernel(const uint32_t* codes, const uint32_t* sel, int64_t n, const uint32_t* dict_off,
                                    uint32_t ndict, uint32_t* len) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
    const uint32_t c = codes[sel ? sel[k] : k];
    len[k] = c < ndict ? dict_off[c + 1] - dict_off[c] : 0u;
  }
}

// output byte p = byte (p % width) of row p / width; booleans as 0 / 1
__global__ void mlb1_pack_fixed_kernel(const uint8_t* src, int shift, int as_bool, const uint32_t* sel, int64_t n,
                                       uint8_t* dst) {
  const int64_t total = n << shift;
  const int64_t mask = ((int64_t)1 << shift) - 1;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < total; p += stride) {
    const int64_t k = p >> shift;
    const int64_t row = sel ? (int64_t)sel[k] : k;
    const uint8_t x = src[(row << shift) + (p & mask)];
    dst[p] = as_bool ? (uint8_t)(x != 0) : x;
  }
}

// one wave per 64-row word; lanes 0-7 store its little-endian bytes
__global__ void mlb1_pack_validity_kernel(const uint8_t* valid, const uint32_t* sel, int64_t n, uint8_t* dst) {
  const int64_t nw = (n + 63) / 64;
  const int64_t waves = (int64_t)gridDim.x * (blockDim.x / 64);
  const int lane = threadIdx.x & 63;
  for (int64_t w = (int64_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64; w < nw; w += waves) {
    const int64_t k = w * 64 + lane;
    const uint64_t m = __ballot(k < n && valid[sel ? sel[k] : k] != 0);
    if (lane < 8) dst[w * 8 + lane] = (uint8_t)(m >> (8 * lane));
  }
}

__global__ void mlb1_pack_strings_kernel(const uint32_t* codes, const uint32_t* sel, int64_t n,
                                         const uint32_t* dict_off, const uint8_t* dict_bytes, uint32_t ndict,
                                         const uint32_t* offs, uint8_t* dst) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
    const uint32_t c = codes[sel ? sel[k] : k];
    if (c >= ndict) continue;
    const uint8_t* src = dict_bytes + dict_off[c];
    const uint32_t len = dict_off[c + 1] - dict_off[c];
    uint8_t* d = dst + offs[k];
    for (uint32_t j = 0; j < len; j++) d[j] = src[j];
  }
}

__global__ void topn_rows_kernel(const uint64_t* words, int nw, int64_t n, int64_t* rows_out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) rows_out[i] = (int64_t)words[(uint64_t)i * (nw + 1) + nw];
}

// =================================================================== launchers
static int grid_for(int64_t n, int block, int max_blocks) {
  int64_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > max_blocks) g = max_blocks;
  return (int)g;
}

void launch_pipeline(const Plan& P, int grid, int block, size_t lds, hipStream_t s) {
  switch (P.sink) {
    case SINK_HASH_AGG: pipeline_kernel<SINK_HASH_AGG><<<grid, block, lds, s>>>(P); break;
    case SINK_DIRECT_AGG: pipeline_kernel<SINK_DIRECT_AGG><<<grid, block, lds, s>>>(P); break;
    case SINK_JOIN_BUILD: pipeline_kernel<SINK_JOIN_BUILD><<<grid, block, lds, s>>>(P); break;
    case SINK_COLLECT: pipeline_kernel<SINK_COLLECT><<<grid, block, lds, s>>>(P); break;
    case SINK_COUNT_PAIRS: pipeline_kernel<SINK_COUNT_PAIRS><<<grid, block, lds, s>>>(P); break;
    case SINK_PROJECT: pipeline_kernel<SINK_PROJECT><<<grid, block, lds, s>>>(P); break;
    case SINK_HASH: pipeline_kernel<SINK_HASH><<<grid, block, lds, s>>>(P); break;
    case SINK_FLAGS: pipeline_kernel<SINK_FLAGS><<<grid, block, lds, s>>>(P); break;
    default: break;
  }
}

void launch_direct_reduce(const uint64_t* partials, int nblocks, int ncells, int nacc, const AccSlotList& accs,
                          uint64_t* out, hipStream_t s) {
  if (ncells <= 0) return;
  direct_tree_kernel<<<ncells, 256, 0, s>>>(partials, nblocks, ncells, nacc, accs, out);
}

template <int MODE>
static void launch_flow_mode(const Plan& P, int grid, int block, size_t lds, hipStream_t s) {
  switch (P.sink) {
    case SINK_HASH_AGG: flow_kernel<SINK_HASH_AGG, MODE><<<grid, block, lds, s>>>(P); break;
    case SINK_DIRECT_AGG: flow_kernel<SINK_DIRECT_AGG, MODE><<<grid, block, lds, s>>>(P); break;
    case SINK_ROWS: flow_kernel<SINK_ROWS, MODE><<<grid, block, lds, s>>>(P); break;
    case SINK_COUNT_PAIRS: flow_kernel<SINK_COUNT_PAIRS, MODE><<<grid, block, lds, s>>>(P); break;
    default: break;
  }
}

void launch_flow(const Plan& P, int grid, int block, size_t lds, hipStream_t s) {
  if (P.fast) launch_flow_mode<FLOW_FAST>(P, grid, block, lds, s);
  else launch_flow_mode<FLOW_GENERIC>(P, grid, block, lds, s);
}

template <int MODE>
static int flow_blocks_mode(int sink, int block, size_t lds) {
  int nb = 0;
  hipError_t e = hipErrorInvalidValue;
  switch (sink) {
    case SINK_HASH_AGG: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, flow_kernel<SINK_HASH_AGG, MODE>, block, lds); break;
    case SINK_DIRECT_AGG: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, flow_kernel<SINK_DIRECT_AGG, MODE>, block, lds); break;
    case SINK_ROWS: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, flow_kernel<SINK_ROWS, MODE>, block, lds); break;
    case SINK_COUNT_PAIRS: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, flow_kernel<SINK_COUNT_PAIRS, MODE>, block, lds); break;
    default: break;
  }
  return (e == hipSuccess && nb > 0) ? nb : 1;
}

int flow_blocks_per_cu(int sink, bool fast, int block, size_t lds) {
  return fast ? flow_blocks_mode<FLOW_FAST>(sink, block, lds) : flow_blocks_mode<FLOW_GENERIC>(sink, block, lds);
}

// Exclusive scan of a workgroup's u64 values (SCAN_T threads); *total = the sum.
__device__ __forceinline__ uint64_t block_excl_scan64(uint64_t v, uint64_t* red, uint64_t* total) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  uint64_t incl = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint64_t t = __shfl_up(incl, o);
    if (lane >= o) incl += t;
  }
  if (lane == 63) red[wv] = incl;
  __syncthreads();
  uint64_t base = 0, all = 0;
  for (int w = 0; w < SCAN_T / 64; w++) {
    if (w < wv) base += red[w];
    all += red[w];
  }
  *total = all;
  __syncthreads();
  return base + incl - v;
}

// Large count scans in three passes over SCAN_B-count blocks, with no buffer
// besides the output: pass 1 parks block b's sum in offsets[b * SCAN_B], pass
// 2 (one workgroup) turns those into the blocks' exclusive bases and writes
// the total at offsets[n], pass 3 writes every block's offsets from its base.
__global__ void __launch_bounds__(SCAN_T) count_scan_reduce_kernel(const uint32_t* counts, int64_t n,
                                                                   uint64_t* offsets) {
  __shared__ uint64_t red[SCAN_T / 64];
  const int64_t base = (int64_t)blockIdx.x * SCAN_B + (int64_t)threadIdx.x * SCAN_E;
  uint64_t c = 0;
#pragma unroll
  for (int e = 0; e < SCAN_E; e++) c += (base + e < n) ? counts[base + e] : 0u;
  uint64_t total;
  block_excl_scan64(c, red, &total);
  if (threadIdx.x == 0) offsets[(int64_t)blockIdx.x * SCAN_B] = total;
}

__global__ void __launch_bounds__(SCAN_T) count_scan_top_kernel(uint64_t* offsets, int64_t n, int64_t nb) {
  __shared__ uint64_t red[SCAN_T / 64];
  uint64_t carry = 0;
  for (int64_t b0 = 0; b0 < nb; b0 += SCAN_T) {
    const int64_t b = b0 + threadIdx.x;
    const uint64_t v = b < nb ? offsets[b * SCAN_B] : 0;
    uint64_t total;
    const uint64_t ex = block_excl_scan64(v, red, &total);
    if (b < nb) offsets[b * SCAN_B] = carry + ex;
    carry += total;
  }
  if (threadIdx.x == 0) offsets[n] = carry;
}

__global__ void __launch_bounds__(SCAN_T) count_scan_down_kernel(const uint32_t* counts, int64_t n,
                                                                 uint64_t* offsets) {
  __shared__ uint64_t red[SCAN_T / 64];
  __shared__ uint64_t block_base;
  const int64_t first = (int64_t)blockIdx.x * SCAN_B;
  if (threadIdx.x == 0) block_base = offsets[first];
  const int64_t base = first + (int64_t)threadIdx.x * SCAN_E;
  uint32_t v[SCAN_E];
  uint64_t c = 0;
#pragma unroll
  for (int e = 0; e < SCAN_E; e++) {
    v[e] = (base + e < n) ? counts[base + e] : 0u;
    c += v[e];
  }
  uint64_t total;
  const uint64_t ex = block_excl_scan64(c, red, &total);  // its barriers order block_base's read first
  uint64_t run = block_base + ex;
#pragma unroll
  for (int e = 0; e < SCAN_E; e++) {
    if (base + e < n) offsets[base + e] = run;
    run += v[e];
  }
}

void launch_scan_counts(const uint32_t* counts, int64_t n, uint64_t* offsets, hipStream_t s) {
  if (n <= 4 * SCAN_B) {  // a few thousand wave counts: one workgroup
    scan_counts_kernel<<<1, 1024, 0, s>>>(counts, n, offsets);
    return;
  }
  const int64_t nb = (n + SCAN_B - 1) / SCAN_B;
  count_scan_reduce_kernel<<<(unsigned)nb, SCAN_T, 0, s>>>(counts, n, offsets);
  count_scan_top_kernel<<<1, SCAN_T, 0, s>>>(offsets, n, nb);
  count_scan_down_kernel<<<(unsigned)nb, SCAN_T, 0, s>>>(counts, n, offsets);
}

void launch_materialize(const Plan& P, int sink, const uint64_t* offsets, int64_t nwaves, hipStream_t s) {
  if (nwaves <= 0) return;
  if (sink == SINK_JOIN_BUILD) materialize_kernel<SINK_JOIN_BUILD><<<(unsigned)nwaves, 256, 0, s>>>(P, offsets);
  else materialize_kernel<SINK_COLLECT><<<(unsigned)nwaves, 256, 0, s>>>(P, offsets);
}

void launch_fill_u64(uint64_t* p, uint64_t v, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  fill_u64_kernel<<<grid_for(n, 256, 8192), 256, 0, s>>>(p, v, n);
}

void launch_gather(const void* src, int width, const int64_t* idx, int64_t n, void* dst, const uint8_t* src_valid,
                   uint8_t* dst_valid, hipStream_t s) {
  if (n <= 0) return;
  gather_kernel<<<grid_for(n, 256, 8192), 256, 0, s>>>(src, width, idx, n, dst, src_valid, dst_valid);
}

void launch_join_insert(const DevJoinTable& T, const uint64_t* keys, const uint8_t* key_null, int64_t n,
                        uint64_t* bloom, hipStream_t s) {
  if (n <= 0) return;
  join_insert_kernel<<<grid_for(n, 256, 8192), 256, 0, s>>>(T, keys, key_null, n, bloom);
}

void launch_join_part_hist(const DevJoinTable& T, const uint64_t* keys, const uint8_t* key_null, int64_t n,
                           int nblocks, int64_t rows_per_block, int nparts, uint32_t* hist, hipStream_t s) {
  join_part_hist_kernel<<<nblocks, 256, (size_t)nparts * 4, s>>>(T, keys, key_null, n, rows_per_block, nparts, hist);
}

void launch_join_part_scatter(const DevJoinTable& T, const uint64_t* keys, const uint8_t* key_null, int64_t n,
                              int nblocks, int64_t rows_per_block, int nparts, const uint32_t* offsets,
                              uint32_t* srow, uint64_t* skeys, hipStream_t s) {
  join_part_scatter_kernel<<<nblocks, 256, (size_t)nparts * 4, s>>>(T, keys, key_null, n, rows_per_block, nparts,
                                                                    offsets, srow, skeys);
}

void launch_join_part_build(const DevJoinTable& T, int nparts, const uint32_t* offsets, int nblocks,
                            const uint32_t* srow, const uint64_t* skeys, uint64_t* bloom, uint32_t* flags,
                            size_t lds, hipStream_t s) {
  join_part_build_kernel<<<nparts, 256, lds, s>>>(T, offsets, nblocks, srow, skeys, bloom, flags);
}

__global__ void strided_u32_kernel(const uint32_t* src, int64_t stride, int64_t n, uint32_t* dst) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[i * stride];
}

void launch_strided_u32(const uint32_t* src, int64_t stride, int64_t n, uint32_t* dst, hipStream_t s) {
  if (n > 0) strided_u32_kernel<<<grid_for(n, 256, 8192), 256, 0, s>>>(src, stride, n, dst);
}

void launch_key_minmax(const uint64_t* keys, int nkw, const uint8_t* key_null, int64_t n, const uint64_t* n_dev,
                       int64_t* partials, int nparts, int64_t* mm, uint32_t* done, uint32_t* clear, hipStream_t s) {
  key_minmax_kernel<<<nparts, 256, 0, s>>>(keys, nkw, key_null, n, n_dev, partials, mm, done, clear);
}

void launch_small_dense(const DensePermute& M, const uint64_t* keys, const uint8_t* key_null, int64_t cap,
                        const uint64_t* count_dev, uint64_t* bits, uint32_t* rank, int64_t* status, hipStream_t s) {
  small_dense_kernel<<<1, 1024, 0, s>>>(M, keys, key_null, cap, count_dev, bits, rank, status);
}
int small_dense_words() { return SMALL_DENSE_WORDS; }

void launch_dense_bloom(const DevJoinTable& T, const uint64_t* keys, const uint8_t* key_null, int64_t n,
                        int64_t kmin, uint64_t range, uint64_t* bloom, hipStream_t s) {
  if (n > 0) dense_bloom_kernel<<<grid_for(n, 256, 8192), 256, 0, s>>>(T, keys, key_null, n, kmin, range, bloom);
}

void launch_dense_mark(const uint64_t* keys, int nkw, const uint8_t* key_null, int64_t n, int64_t kmin,
                       uint64_t range, uint64_t* bits, uint32_t* flag, hipStream_t s) {
  if (n > 0)
    dense_mark_kernel<<<grid_for(n, 256, 8192), 256, 0, s>>>(keys, nkw, key_null, n, kmin, range, bits, flag);
}

static void scan_launch(const void* in, int popc, int64_t n, uint32_t* out, uint32_t* block_sums, hipStream_t s);

// Rows of an offset-indexed dense table (position p: a key offset below
// `range`, or range + i for the i-th NULL-key row) -> their row numbers in key
// order: the rank of the key's bit, the NULL-key rows after all keyed ones.
__device__ __forceinline__ uint64_t dense_compact_row(const uint64_t* bits, const uint32_t* rank, uint64_t range,
                                                      uint64_t p) {
  if (p >= range) return (uint64_t)rank[(range + 63) >> 6] + (p - range);
  return (uint64_t)rank[p >> 6] + (uint64_t)__popcll(bits[p >> 6] & ((1ull << (p & 63)) - 1ull));
}

__global__ void dense_translate_kernel(int64_t* rows, int64_t stride, int64_t n, const uint64_t* bits,
                                       const uint32_t* rank, uint64_t range) {
  const int64_t step = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += step)
    rows[i * stride] = (int64_t)dense_compact_row(bits, rank, range, (uint64_t)rows[i * stride]);
}

__global__ void dense_compact_kernel(const uint8_t* src, int width, const uint8_t* src_valid, int64_t npos,
                                     const uint64_t* bits, const uint32_t* rank, uint64_t range, uint8_t* dst,
                                     uint8_t* dst_valid) {
  const int64_t step = (int64_t)gridDim.x * blockDim.x;
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < npos; p += step) {
    if ((uint64_t)p < range && !((bits[p >> 6] >> (p & 63)) & 1ull)) continue;  // no row at this offset
    const uint64_t q = dense_compact_row(bits, rank, range, (uint64_t)p);
    for (int b = 0; b < width; b++) dst[q * width + b] = src[(uint64_t)p * width + b];
    if (dst_valid) dst_valid[q] = src_valid ? src_valid[p] : 1;
  }
}

void launch_dense_translate(int64_t* rows, int64_t stride, int64_t n, const uint64_t* bits, const uint32_t* rank,
                            uint64_t range, hipStream_t s) {
  if (n > 0) dense_translate_kernel<<<grid_for(n, 256, 8192), 256, 0, s>>>(rows, stride, n, bits, rank, range);
}

void launch_dense_compact(const void* src, int width, const uint8_t* src_valid, int64_t npos, const uint64_t* bits,
                          const uint32_t* rank, uint64_t range, void* dst, uint8_t* dst_valid, hipStream_t s) {
  if (npos > 0)
    dense_compact_kernel<<<grid_for(npos, 256, 8192), 256, 0, s>>>((const uint8_t*)src, width, src_valid, npos, bits,
                                                                   rank, range, (uint8_t*)dst, dst_valid);
}

void launch_dense_rank(const uint64_t* bits, int64_t nw, uint32_t* rank, uint32_t* block_sums, hipStream_t s) {
  scan_launch(bits, 1, nw, rank, block_sums, s);
}

__global__ void dense_key_fill_kernel(uint8_t* dst, int width, int64_t held, uint64_t range, int64_t kmin) {
  const int64_t step = (int64_t)gridDim.x * blockDim.x;
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < held; p += step) {
    const int64_t v = (uint64_t)p < range ? kmin + p : 0;
    if (width == 8) ((int64_t*)dst)[p] = v;
    else ((int32_t*)dst)[p] = (int32_t)v;
  }
}

void launch_dense_key_fill(void* dst, int width, int64_t held, uint64_t range, int64_t kmin, hipStream_t s) {
  if (held > 0)
    dense_key_fill_kernel<<<grid_for(held, 256, 8192), 256, 0, s>>>((uint8_t*)dst, width, held, range, kmin);
}

void launch_dense_place(const DensePermute& M, const uint64_t* keys, const uint8_t* key_null, int64_t n, int64_t kmin,
                        uint64_t range, uint64_t* bits, uint32_t* flag, uint32_t* null_next, hipStream_t s) {
  if (n > 0)
    dense_place_kernel<<<grid_for(n, 256, 8192), 256, 0, s>>>(M, keys, key_null, n, kmin, range, bits, flag, null_next);
}

void launch_key_range_reduce(const int64_t* partials, int nparts, int64_t n, const uint64_t* n_dev, int64_t* mm,
                             uint32_t* clear, hipStream_t s) {
  key_range_reduce_kernel<<<1, 256, 0, s>>>(partials, nparts, n, n_dev, mm, clear);
}

void launch_dense_permute(const DensePermute& M, const uint64_t* keys, const uint8_t* key_null, int64_t n,
                          int64_t kmin, uint64_t range, const uint64_t* bits, const uint32_t* rank,
                          const uint32_t* present, uint32_t* null_next, hipStream_t s) {
  if (n > 0)
    dense_permute_kernel<<<grid_for(n, 256, 8192), 256, 0, s>>>(M, keys, key_null, n, kmin, range, bits, rank,
                                                                 present, null_next);
}

void launch_agg_part_hist(const AggRecords& R, int64_t n, uint64_t mask, int part_shift, int nblocks,
                          int64_t rows_per_block, int nparts, uint32_t* hist, hipStream_t s) {
  agg_part_hist_kernel<<<nblocks, PART_PASS_THREADS, (size_t)nparts * 4, s>>>(R, n, mask, part_shift, rows_per_block,
                                                                              nparts, hist);
}

void launch_agg_part_scatter(const AggRecords& R, int64_t n, uint64_t mask, int part_shift, int nblocks,
                             int64_t rows_per_block, int nparts, const uint32_t* offsets, uint64_t* rkeys,
                             uint64_t* rargs, uint32_t* rvalid, hipStream_t s) {
  agg_part_scatter_kernel<<<nblocks, PART_PASS_THREADS, (size_t)nparts * 4, s>>>(R, n, mask, part_shift,
                                                                                 rows_per_block, nparts, offsets,
                                                                                 rkeys, rargs, rvalid);
}

void launch_agg_part_chunks(const PartAggPlan& A, const PartChunk* chunks, int nchunks, const uint64_t* rkeys,
                            const uint64_t* rargs, const uint32_t* rvalid, int64_t n, uint64_t* partial, hipStream_t s) {
  if (nchunks <= 0) return;
  agg_part_chunk_kernel<<<nchunks, 256, agg_part_lds(A), s>>>(A, chunks, rkeys, rargs, rvalid, n, partial);
}

void launch_agg_part_merge(const PartAggPlan& A, const uint32_t* parts, const uint32_t* first, int nparts,
                           const uint64_t* partial, hipStream_t s) {
  if (nparts <= 0) return;
  agg_part_merge_kernel<<<nparts, MERGE_THREADS, agg_part_lds(A), s>>>(A, parts, first, partial);
}

int64_t scan_u32_blocks(int64_t n) { return (n + SCAN_B - 1) / SCAN_B; }

static void scan_launch(const void* in, int popc, int64_t n, uint32_t* out, uint32_t* block_sums, hipStream_t s) {
  const int64_t nb = scan_u32_blocks(n);
  if (nb == 0) {
    scan_u32_top_kernel<<<1, SCAN_T, 0, s>>>(out, 0);  // out[0] = 0
    return;
  }
  scan_u32_reduce_kernel<<<(unsigned)nb, SCAN_T, 0, s>>>(in, popc, n, block_sums);
  scan_u32_top_kernel<<<1, SCAN_T, 0, s>>>(block_sums, nb);
  scan_u32_down_kernel<<<(unsigned)nb, SCAN_T, 0, s>>>(in, popc, n, block_sums, out);
}

void launch_scan_u32(const uint32_t* in, int64_t n, uint32_t* out, uint32_t* block_sums, hipStream_t s) {
  scan_launch(in, 0, n, out, block_sums, s);
}

void launch_join_unique(const DevJoinTable& T, uint32_t* dup_flag, hipStream_t s) {
  join_unique_kernel<<<grid_for((int64_t)T.mask + 1, 256, 8192), 256, 0, s>>>(T, dup_flag);
}

void launch_agg_rehash(const uint64_t* old_slots, uint64_t old_cap, int nkw, uint64_t* new_slots, uint64_t new_mask,
                       const AccMove& mv, hipStream_t s) {
  agg_rehash_kernel<<<grid_for((int64_t)old_cap, 256, 8192), 256, 0, s>>>(old_slots, old_cap, nkw, new_slots,
                                                                          new_mask, mv);
}

void launch_slot_flags(const uint64_t* slots, uint64_t cap, int nkw, uint8_t* flags, hipStream_t s) {
  slot_flags_kernel<<<grid_for((int64_t)cap, 256, 8192), 256, 0, s>>>(slots, cap, nkw, flags);
}

void launch_agg_emit(const uint64_t* slots, int nkw, const int64_t* idx, int64_t n, const EmitPlan& E,
                     hipStream_t s) {
  if (n <= 0) return;
  agg_emit_kernel<<<grid_for(n, 256, 8192), 256, 0, s>>>(slots, nkw, idx, n, E);
}

int64_t compact_blocks(int64_t n) { return (n + CB - 1) / CB; }

void launch_compact(const uint8_t* flags, int64_t n, uint32_t* block_counts, uint64_t* offsets, const uint32_t* sel,
                    void* out, int out_u32, hipStream_t s) {
  const int64_t nb = compact_blocks(n);
  if (nb == 0) {
    fill_u64_kernel<<<1, 64, 0, s>>>(offsets, 0, 1);
    return;
  }
  flag_counts_kernel<<<(unsigned)nb, CT, 0, s>>>(flags, n, block_counts);
  scan_counts_kernel<<<1, 1024, 0, s>>>(block_counts, nb, offsets);
  flag_scatter_kernel<<<(unsigned)nb, CT, 0, s>>>(flags, n, offsets, sel, out, out_u32);
}

void launch_sort_keys(const Plan& P, const SortKeys& K, int64_t n, uint64_t* words, hipStream_t s) {
  if (n <= 0) return;
  sort_keys_kernel<<<grid_for(n, 256, 8192), 256, 0, s>>>(P, K, n, words);
}

void launch_topn_tiles(const uint64_t* in, int64_t n, int nw, int keep, uint64_t* out, hipStream_t s) {
  const int64_t tiles = (n + TOPN_TILE - 1) / TOPN_TILE;
  const size_t lds = (size_t)TOPN_TILE * (nw + 1) * 8;
  switch (nw + 1) {
    case 2: topn_tile_kernel<2><<<(unsigned)tiles, 256, lds, s>>>(in, n, keep, out); break;
    case 3: topn_tile_kernel<3><<<(unsigned)tiles, 256, lds, s>>>(in, n, keep, out); break;
    case 4: topn_tile_kernel<4><<<(unsigned)tiles, 256, lds, s>>>(in, n, keep, out); break;
    case 5: topn_tile_kernel<5><<<(unsigned)tiles, 256, lds, s>>>(in, n, keep, out); break;
    case 6: topn_tile_kernel<6><<<(unsigned)tiles, 256, lds, s>>>(in, n, keep, out); break;
    default: break;  // the caller refuses wider keys
  }
}

void launch_topn_rank(const uint64_t* in, int64_t n, int nw, int keep, uint64_t* out, hipStream_t s) {
  const unsigned tiles = (unsigned)((n + TOPN_RANK_TILE - 1) / TOPN_RANK_TILE);
  switch (nw + 1) {
    case 2: topn_rank_kernel<2><<<tiles, TOPN_RANK_TILE * TOPN_RANK_SPLIT, 0, s>>>(in, n, keep, out); break;
    case 3: topn_rank_kernel<3><<<tiles, TOPN_RANK_TILE * TOPN_RANK_SPLIT, 0, s>>>(in, n, keep, out); break;
    case 4: topn_rank_kernel<4><<<tiles, TOPN_RANK_TILE * TOPN_RANK_SPLIT, 0, s>>>(in, n, keep, out); break;
    case 5: topn_rank_kernel<5><<<tiles, TOPN_RANK_TILE * TOPN_RANK_SPLIT, 0, s>>>(in, n, keep, out); break;
    case 6: topn_rank_kernel<6><<<tiles, TOPN_RANK_TILE * TOPN_RANK_SPLIT, 0, s>>>(in, n, keep, out); break;
    default: break;  // the caller refuses wider keys
  }
}

int64_t sort_padded_entries(int64_t n) { return std::max<int64_t>(1, (n + TOPN_TILE - 1) / TOPN_TILE) * TOPN_TILE; }

int64_t sort_split_entries(int64_t n) { return sort_padded_entries(n) / MERGE_OUT + 1; }

uint64_t* launch_sort_words(const uint64_t* words, int64_t n, int nw, uint64_t* a, uint64_t* b, int64_t* split,
                            hipStream_t s) {
  const int W = nw + 1;
  const int64_t N = sort_padded_entries(n);
  const int64_t nblocks = N / MERGE_OUT;
  // sorted tiles (the entries past n are all-ones words, which sort last)
  launch_topn_tiles(words, n, nw, TOPN_TILE, a, s);
  uint64_t* cur = a;
  uint64_t* nxt = b;
  for (int64_t L = TOPN_TILE; L < N; L *= 2) {
    const unsigned gs = (unsigned)((nblocks + 255) / 256), gp = (unsigned)nblocks;
    const size_t lds = (size_t)MERGE_OUT * W * 8;
    switch (W) {
#define ML_MERGE(w)                                                                 \
  case w:                                                                           \
    merge_split_kernel<w><<<gs, 256, 0, s>>>(cur, N, L, nblocks, split);            \
    merge_pass_kernel<w><<<gp, 256, lds, s>>>(cur, nxt, N, L, split);               \
    break;
      ML_MERGE(2) ML_MERGE(3) ML_MERGE(4) ML_MERGE(5) ML_MERGE(6)
#undef ML_MERGE
      default: break;  // the caller refuses wider keys
    }
    std::swap(cur, nxt);
  }
  return cur;
}

void launch_null_count(const uint8_t* valid, const uint32_t* sel, int64_t n, uint64_t* out, hipStream_t s) {
  if (n <= 0) return;
  null_count_kernel<<<grid_for(n, 256, 2048), 256, 0, s>>>(valid, sel, n, out);
}

void launch_mlb1_str_len(const uint32_t* codes, const uint32_t* sel, int64_t n, const uint32_t* dict_off,
                         uint32_t ndict, uint32_t* len, hipStream_t s) {
  if (n <= 0) return;
  mlb1_str_len_kernel<<<grid_for(n, 256, 8192), 256, 0, s>>>(codes, sel, n, dict_off, ndict, len);
}

void launch_mlb1_pack_fixed(const void* src, int width, bool as_bool, const uint32_t* sel, int64_t n, uint8_t* dst,
                            hipStream_t s) {
  if (n <= 0) return;
  const int shift = width == 8 ? 3 : (width == 4 ? 2 : 0);
  mlb1_pack_fixed_kernel<<<grid_for(n << shift, 256, 8192), 256, 0, s>>>((const uint8_t*)src, shift, as_bool ? 1 : 0,
                                                                         sel, n, dst);
}

void launch_mlb1_pack_validity(const uint8_t* valid, const uint32_t* sel, int64_t n, uint8_t* dst, hipStream_t s) {
  if (n <= 0) return;
  mlb1_pack_validity_kernel<<<grid_for((n + 63) / 64, 4, 8192), 256, 0, s>>>(valid, sel, n, dst);
}

void launch_mlb1_pack_strings(const uint32_t* codes, const uint32_t* sel, int64_t n, const uint32_t* dict_off,
                              const uint8_t* dict_bytes, uint32_t ndict, const uint32_t* offs, uint8_t* dst,
                              hipStream_t s) {
  if (n <= 0) return;
  mlb1_pack_strings_kernel<<<grid_for(n, 256, 8192), 256, 0, s>>>(codes, sel, n, dict_off, dict_bytes, ndict, offs,
                                                                   dst);
}

void launch_bits_to_bytes(const uint64_t* bits, int64_t n, uint8_t* out, hipStream_t s) {
  if (n <= 0) return;
  bits_to_bytes_kernel<<<grid_for(n, 256, 8192), 256, 0, s>>>(bits, n, out);
}

void launch_bytes_to_bits(const uint8_t* valid, int64_t n, uint64_t* bits, hipStream_t s) {
  if (n <= 0) return;
  bytes_to_bits_kernel<<<grid_for((n + 63) / 64, 4, 8192), 256, 0, s>>>(valid, n, bits);
}

void launch_topn_rows(const uint64_t* words, int nw, int64_t n, int64_t* rows_out, hipStream_t s) {
  if (n <= 0) return;
  topn_rows_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(words, nw, n, rows_out);
}

}  // namespace mlg
