"""Write tests/golden/hash_vectors.json: known answers for MiniLake's folded-multiply hash.

The values are computed here with Python's arbitrary-precision integers
directly from the published definitions (crates/minilake-hashtable/src/hash.rs:12-55):
  fold_mul(a, b) = low64(a*b) XOR high64(a*b)
  hash_u64(x)    = fold_mul(x XOR SEED, MUL)
  combine(h, v)  = fold_mul(h XOR v, MUL2)
  hash_bytes(b)  = 8-byte little-endian words folded with MUL, zero-padded tail with MUL2,
                   starting from SEED XOR (len * MUL mod 2^64)
and the hash_columns rule for key columns (crates/minilake-exec/src/kernels/hash.rs:16-79):
i32/Date sign-extended, f64 +-0 -> 0 else bit pattern, NULL -> NULL_HASH.
"""

import json
import os
import struct

M64 = (1 << 64) - 1
SEED = 0x243F6A8885A308D3
MUL = 0x9E3779B97F4A7C15
MUL2 = 0xBF58476D1CE4E5B9
NULL_HASH = 0x7C159E3779B97F4A


def fold_mul(a, b):
    r = (a & M64) * (b & M64)
    return (r & M64) ^ (r >> 64)


def hash_u64(x):
    return fold_mul((x & M64) ^ SEED, MUL)


def combine(h, v):
    return fold_mul(h ^ v, MUL2)


def hash_bytes(b: bytes):
    h = SEED ^ ((len(b) * MUL) & M64)
    full = len(b) // 8
    for i in range(full):
        h = fold_mul(h ^ int.from_bytes(b[8 * i:8 * i + 8], "little"), MUL)
    tail = b[8 * full:]
    if tail:
        h = fold_mul(h ^ int.from_bytes(tail + b"\0" * (8 - len(tail)), "little"), MUL2)
    return h


def f64_bits(x):
    return 0 if x == 0.0 else struct.unpack("<Q", struct.pack("<d", x))[0]


def main():
    u64s = [0, 1, 2, 3, 42, 255, 256, 65535, 1 << 31, (1 << 32) - 1, 1 << 32, (1 << 63) - 1, 1 << 63, M64,
            0x243F6A8885A308D3, 0x9E3779B97F4A7C15, 123456789012345678, 0xDEADBEEFCAFEBABE]
    i32s = [0, 1, -1, 7, -7, 2147483647, -2147483648, 10471, 9204]
    f64s = [0.0, -0.0, 1.0, -1.0, 0.5, 0.1, 900.0, 105000.0, 1e300, -2.5e-10]
    strs = ["", "\0", "a", "A", "N", "R", "F", "O", "abc", "abd", "BUILDING", "MACHINERY", "12345678", "123456789",
            "UNITED STATES", "MIDDLE EAST", "a much longer string with more than sixteen bytes"]
    out = {
        "constants": {"SEED": SEED, "MUL": MUL, "MUL2": MUL2, "NULL_HASH": NULL_HASH},
        "fold_mul": [[a, b, fold_mul(a, b)] for a in u64s[:8] for b in u64s[8:]],
        "hash_u64": [[x, hash_u64(x)] for x in u64s],
        "combine": [[h, v, combine(h, v)] for h in u64s[:6] for v in u64s[6:12]],
        "hash_bytes": [[s, hash_bytes(s.encode()), len(s.encode())] for s in strs],
        # hash_columns over one column per type, seeded with SEED
        "column_i32": [[x, combine(SEED, hash_u64(x & M64))] for x in i32s],
        "column_i64": [[x - (1 << 64) if x >= 1 << 63 else x, combine(SEED, hash_u64(x))] for x in u64s],
        "column_f64": [[x, combine(SEED, hash_u64(f64_bits(x)))] for x in f64s],
        "column_bool": [[b, combine(SEED, hash_u64(int(b)))] for b in (False, True)],
        "column_str": [[s, combine(SEED, hash_bytes(s.encode()))] for s in strs],
        "null_single": combine(SEED, NULL_HASH),
        # a three-column key: (i64, date, i32) as in the Q3 GROUP BY
        "three_keys": [[a, b, c, combine(combine(combine(SEED, hash_u64(a & M64)), hash_u64(b & M64)), hash_u64(c & M64))]
                       for a, b, c in [(1, 9000, 0), (6000000, 8035, 4), (-5, -1, 2147483647)]],
    }
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "hash_vectors.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(path)


if __name__ == "__main__":
    main()
