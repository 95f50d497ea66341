// Integer bit helpers shared by the WZP kernels. Host+device so the CPU test
// harness (tests/test_bits_host.cpp) checks them against brute force.
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define WZB_HD __host__ __device__ __forceinline__
#else
#define WZB_HD inline
#endif

namespace wz {

WZB_HD uint32_t popc32(uint32_t v) { return uint32_t(__builtin_popcount(v)); }
WZB_HD uint32_t popc64(uint64_t v) { return uint32_t(__builtin_popcountll(v)); }
#if defined(__clang__)
WZB_HD uint32_t bitrev32(uint32_t v) { return __builtin_bitreverse32(v); }
#else
WZB_HD uint32_t bitrev32(uint32_t v) {
    uint32_t r = 0;
    for (int i = 0; i < 32; ++i) r |= ((v >> i) & 1u) << (31 - i);
    return r;
}
#endif
WZB_HD uint32_t bitrev16(uint32_t v) { return bitrev32(v) >> 16; }

// In-place transpose of a 16x16 bit matrix held as 16 rows of 16 bits:
// afterwards row c bit r holds what row r bit c held (Hacker's Delight
// block-swap, four rounds of masked exchanges).
WZB_HD void transpose16(uint32_t (&a)[16]) {
#define WZ_SWAP(J, M)                                                   \
    _Pragma("unroll") for (int k = 0; k < 16; ++k)                                        \
        if (!(k & (J))) {                                               \
            const uint32_t t = ((a[k] >> (J)) ^ a[k + (J)]) & (M);      \
            a[k + (J)] ^= t;                                            \
            a[k] ^= t << (J);                                           \
        }
    WZ_SWAP(8, 0x00FFu)
    WZ_SWAP(4, 0x0F0Fu)
    WZ_SWAP(2, 0x3333u)
    WZ_SWAP(1, 0x5555u)
#undef WZ_SWAP
}

// Two independent 16x16 transposes, one per 16-bit half of each word: the
// same masked exchanges with the masks repeated in both halves (no exchange
// crosses bit 16, so the halves never mix).
WZB_HD void transpose16x2(uint32_t (&a)[16]) {
#define WZ_SWAP(J, M)                                                   \
    _Pragma("unroll") for (int k = 0; k < 16; ++k)                                        \
        if (!(k & (J))) {                                               \
            const uint32_t t = ((a[k] >> (J)) ^ a[k + (J)]) & (M);      \
            a[k + (J)] ^= t;                                            \
            a[k] ^= t << (J);                                           \
        }
    WZ_SWAP(8, 0x00FF00FFu)
    WZ_SWAP(4, 0x0F0F0F0Fu)
    WZ_SWAP(2, 0x33333333u)
    WZ_SWAP(1, 0x55555555u)
#undef WZ_SWAP
}

// BIT_2 of one 16-magnitude block: plane p bit (15-j) = magnitude j bit (15-p)
// (the LC BIT_2 orientation of cpu/src/wzp_cpu.cpp:146-193).
// In: m[j] magnitudes. Out: m[p] planes.
WZB_HD void bit2_forward(uint32_t (&m)[16]) {
    uint32_t a[16];
    for (int q = 0; q < 16; ++q) a[q] = bitrev16(m[15 - q]);
    transpose16(a);
    for (int p = 0; p < 16; ++p) m[p] = a[p];
}
// Inverse of bit2_forward. In: planes. Out: magnitudes.
WZB_HD void bit2_inverse(uint32_t (&m)[16]) {
    uint32_t a[16];
    for (int p = 0; p < 16; ++p) a[p] = m[p];
    transpose16(a);
    for (int q = 0; q < 16; ++q) m[15 - q] = bitrev16(a[q]);
}

// BIT_2 of two consecutive blocks m[0..15] and m[16..31] (16-bit magnitudes):
// the low half of planes[p] is plane p of the first block, the high half
// plane p of the second. One 32-bit reversal serves both halves:
// bitrev32(hi << 16 | lo) = bitrev16(hi) | bitrev16(lo) << 16.
WZB_HD void bit2_forward_pair(const uint32_t (&m)[32], uint32_t (&planes)[16]) {
#pragma unroll
    for (int q = 0; q < 16; ++q) planes[q] = bitrev32((m[15 - q] << 16) | m[31 - q]);
    transpose16x2(planes);
}
// Inverse of bit2_forward_pair.
WZB_HD void bit2_inverse_pair(const uint32_t (&planes)[16], uint32_t (&m)[32]) {
    uint32_t a[16];
#pragma unroll
    for (int p = 0; p < 16; ++p) a[p] = planes[p];
    transpose16x2(a);
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const uint32_t r = bitrev32(a[q]);
        m[15 - q] = r >> 16;
        m[31 - q] = r & 0xFFFFu;
    }
}

// Gather the bits of v selected by mask into the low end, in order (x86
// pext), branch-free in five rounds: round i moves every selected bit right
// by 2^i when the count of unselected bits below it has bit i set (the
// "compress" method of Hacker's Delight, section 7-4).
WZB_HD uint32_t compress_bits(uint32_t v, uint32_t mask) {
    uint32_t x = v & mask, m = mask, mk = ~mask << 1;
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        uint32_t mp = mk ^ (mk << 1);  // parity of the zeros below: prefix XOR
        mp ^= mp << 2;
        mp ^= mp << 4;
        mp ^= mp << 8;
        mp ^= mp << 16;
        const uint32_t mv = mp & m;  // selected bits that move by 2^i this round
        m = (m ^ mv) | (mv >> (1 << i));
        const uint32_t t = x & mv;
        x = (x ^ t) | (t >> (1 << i));
        mk &= ~mp;
    }
    return x;
}
// Scatter the low bits of v to the positions selected by mask (x86 pdep):
// the compress rounds run forward to record which bits move, then the moves
// are undone in reverse order (Hacker's Delight, section 7-5).
WZB_HD uint32_t expand_bits(uint32_t v, uint32_t mask) {
    uint32_t m = mask, mk = ~mask << 1, moved[5];
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        uint32_t mp = mk ^ (mk << 1);
        mp ^= mp << 2;
        mp ^= mp << 4;
        mp ^= mp << 8;
        mp ^= mp << 16;
        const uint32_t mv = mp & m;
        moved[i] = mv;
        m = (m ^ mv) | (mv >> (1 << i));
        mk &= ~mp;
    }
    uint32_t x = v;
#pragma unroll
    for (int i = 4; i >= 0; --i) {
        const uint32_t mv = moved[i], t = x << (1 << i);
        x = (x & ~mv) | (t & mv);
    }
    return x & mask;
}

// Per-byte nonzero flags of a 32-bit word packed into 4 bits (byte i -> bit i).
WZB_HD uint32_t nonzero_bytes4(uint32_t w) {
    const uint32_t hi = ((w & 0x7F7F7F7Fu) + 0x7F7F7F7Fu | w) & 0x80808080u;
    return ((hi >> 7) | (hi >> 14) | (hi >> 21) | (hi >> 28)) & 0xFu;
}
// Per-byte nonzero flags of a 64-bit word packed into 8 bits.
WZB_HD uint32_t nonzero_bytes8(uint64_t w) {
    return nonzero_bytes4(uint32_t(w)) | (nonzero_bytes4(uint32_t(w >> 32)) << 4);
}
// RZE change flags for 8 consecutive bitmap bytes: bit k set iff byte k of w
// differs from byte k-1 (byte -1 is `prev`) (cpu/src/wzp_cpu.cpp:566-577).
WZB_HD uint32_t change_flags8(uint64_t w, uint32_t prev) {
    const uint64_t shifted = (w << 8) | uint64_t(prev & 0xFFu);
    return nonzero_bytes8(w ^ shifted);
}

}  // namespace wz
