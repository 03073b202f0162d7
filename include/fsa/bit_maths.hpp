// bit_maths.hpp -- drop-in for the reference's src/bit_maths.hpp host helpers.
// The amplitude-index arithmetic itself runs on the GPU (device functions in
// the library); these remain for callers that build masks or index lists.
#ifndef BIT_MATHS_HPP
#define BIT_MATHS_HPP

#include "types.hpp"

#define INLINE static inline __attribute__((always_inline))

INLINE Index powerOf2(Nat e) { return 1ULL << e; }
INLINE bool isPowerOf2(Index n) { return n > 0 && (n & (n - 1U)) == 0; }
INLINE Nat getBit(Index n, Nat b) { return (n >> b) & 1U; }
INLINE Index flipBit(Index n, Nat b) { return n ^ (1ULL << b); }
INLINE Index insertBit(Index n, Nat b, Nat v) {
    return ((n >> b) << (b + 1)) | ((Index)v << b) | (n & ((1ULL << b) - 1));
}
INLINE Index insertBits(Index n, const NatArray& ascending, Nat v) {
    for (Nat b : ascending) n = insertBit(n, b, v);
    return n;
}
// bit i of value -> position bitIndices[i] (the evident intent of
// bit_maths.hpp:62-78; the shipped setBit drops the other bits)
INLINE Index setBit(Index n, Nat b, Nat v) { return (n & ~(1ULL << b)) | ((Index)v << b); }
INLINE Index setBits(Index n, const NatArray& bitIndices, Index value) {
    for (Nat i = 0; i < bitIndices.size(); i++) n = setBit(n, bitIndices[i], getBit(value, i));
    return n;
}
INLINE Nat getBitMaskParity(Index m) { return (Nat)__builtin_parityll(m); }
static inline Index getBitMask(const NatArray& bits) {
    Index m = 0;
    for (Nat b : bits) m = flipBit(m, b);
    return m;
}
static inline Nat logBase2(Index p) {
    Nat e = 0;
    while (!getBit(p, 0)) { e++; p >>= 1; }
    return e;
}

#endif  // BIT_MATHS_HPP
