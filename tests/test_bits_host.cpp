// Host check of the WZP bit helpers in waltz_amd/csrc/bits.hpp against brute
// force definitions. Built and run by tests/test_bits.py (CPU only).
#include "../waltz_amd/csrc/bits.hpp"

#include <cstdio>
#include <cstdlib>

using namespace wz;

static uint64_t state = 88172645463325252ull;
static uint32_t rnd() {
    state ^= state << 13;
    state ^= state >> 7;
    state ^= state << 17;
    return uint32_t(state);
}

int main() {
    int bad = 0;
    for (int iter = 0; iter < 20000; ++iter) {
        uint32_t m[16], p[16];
        for (int j = 0; j < 16; ++j) {
            m[j] = rnd() & 0xFFFFu;
            if (iter % 3 == 0) m[j] &= (1u << (rnd() % 17)) - 1u;
            p[j] = m[j];
        }
        bit2_forward(p);
        for (int pl = 0; pl < 16; ++pl) {
            uint32_t want = 0;
            for (int j = 0; j < 16; ++j) want |= ((m[j] >> (15 - pl)) & 1u) << (15 - j);
            if (p[pl] != want) ++bad;
        }
        bit2_inverse(p);
        for (int j = 0; j < 16; ++j) if (p[j] != m[j]) ++bad;

        const uint32_t v = rnd(), mask = rnd() & rnd();
        uint32_t c = 0, k = 0;
        for (int i = 0; i < 32; ++i)
            if ((mask >> i) & 1u) c |= ((v >> i) & 1u) << k++;
        if (compress_bits(v, mask) != c) ++bad;
        if (expand_bits(c, mask) != (v & mask)) ++bad;

        uint64_t w = (uint64_t(rnd()) << 32) | rnd();
        if (iter & 1) w &= 0xFF00FF0000FF00FFull;
        uint32_t nz = 0, ch = 0;
        const uint32_t prev = rnd() & 0xFFu;
        for (int b = 0; b < 8; ++b) {
            const uint32_t byte = uint32_t(w >> (8 * b)) & 0xFFu;
            const uint32_t before = b ? uint32_t(w >> (8 * (b - 1))) & 0xFFu : prev;
            nz |= uint32_t(byte != 0) << b;
            ch |= uint32_t(byte != before) << b;
        }
        if (nonzero_bytes8(w) != nz) ++bad;
        if (change_flags8(w, prev) != ch) ++bad;
    }
    std::printf("bits: %d mismatches\n", bad);
    return bad ? 1 : 0;
}
