// Host check of the WZP bit helpers in waltz_amd/csrc/bits.hpp against brute
// force definitions. Built and run by tests/test_bits.py (CPU only).
#include "../waltz_amd/csrc/bits.hpp"
#include "../waltz_amd/csrc/pairbrick.hpp"

#include <cstdio>
#include <cstdlib>
#include <vector>

using namespace wz;

static uint64_t state = 88172645463325252ull;
static uint32_t rnd() {
    state ^= state << 13;
    state ^= state >> 7;
    state ^= state << 17;
    return uint32_t(state);
}

// Recursive ceil-half octant rank of (x, y, z) in an nx x ny x nz brick
// (cpu/src/codec.cpp:96-107), restated plainly.
static uint32_t octant_rank(uint32_t x, uint32_t y, uint32_t z, uint32_t nx, uint32_t ny, uint32_t nz) {
    uint32_t rank = 0;
    while (nx > 1 || ny > 1 || nz > 1) {
        const uint32_t hx = nx - nx / 2, hy = ny - ny / 2, hz = nz - nz / 2;
        const uint32_t sz = z >= hz ? nz / 2 : hz, sy = y >= hy ? ny / 2 : hy, sx = x >= hx ? nx / 2 : hx;
        if (z >= hz) { rank += hz * ny * nx; z -= hz; }
        if (y >= hy) { rank += sz * hy * nx; y -= hy; }
        if (x >= hx) { rank += sz * sy * hx; x -= hx; }
        nx = sx;
        ny = sy;
        nz = sz;
    }
    return rank;
}

// Every sample of a field lands on the magnitude position the brick layout
// gives it (base of brick + octant rank) through the paired-brick map.
static int check_pair_bricks(uint32_t nx, uint32_t ny, uint32_t nz) {
    std::vector<uint32_t> pos(size_t(nx) * ny * nz, 0xFFFFFFFFu);
    for (uint32_t z = 0; z < nz; ++z)
        for (uint32_t y = 0; y < ny; ++y)
            for (uint32_t x = 0; x < nx; ++x) {
                const uint32_t x0 = x / 16 * 16, y0 = y / 16 * 16, z0 = z / 16 * 16;
                const uint32_t base = z0 * nx * ny + y0 * nx * 16 + x0 * 16 * 16;
                pos[(size_t(z) * ny + y) * nx + x] = base + octant_rank(x - x0, y - y0, z - z0, 16, 16, 16);
            }
    int bad = 0;
    std::vector<uint8_t> seen(pos.size(), 0);
    const uint32_t chunks = nx * ny * nz / 8192;
    for (uint32_t c = 0; c < chunks; ++c)
        for (uint32_t t = 0; t < 256; ++t) {
            uint32_t X, Y, Z;
            pair_subblock(nx, ny, c, t, X, Y, Z);
            for (uint32_t dz = 0; dz < 2; ++dz)
                for (uint32_t dy = 0; dy < 4; ++dy)
                    for (uint32_t dx = 0; dx < 4; ++dx) {
                        const size_t at = (size_t(Z + dz) * ny + Y + dy) * nx + X + dx;
                        if (X + dx >= nx || Y + dy >= ny || Z + dz >= nz) { ++bad; continue; }
                        if (pos[at] != c * 8192 + 32 * t + subblock_rank(dx, dy, dz)) ++bad;
                        ++seen[at];
                    }
        }
    for (uint8_t v : seen) bad += v != 1;
    return bad;
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

        {  // the two-block form against two one-block transforms
            uint32_t m2[32], a[16], b[16], planes[16], back[32];
            for (int j = 0; j < 32; ++j) {
                m2[j] = rnd() & 0xFFFFu;
                if (iter % 5 == 0) m2[j] &= (1u << (rnd() % 17)) - 1u;
            }
            for (int j = 0; j < 16; ++j) {
                a[j] = m2[j];
                b[j] = m2[16 + j];
            }
            bit2_forward(a);
            bit2_forward(b);
            bit2_forward_pair(m2, planes);
            for (int pl = 0; pl < 16; ++pl)
                if (planes[pl] != (a[pl] | (b[pl] << 16))) ++bad;
            bit2_inverse_pair(planes, back);
            for (int j = 0; j < 32; ++j)
                if (back[j] != m2[j]) ++bad;
        }

        const uint32_t v = rnd();
        const uint32_t mask = iter % 4 == 0 ? rnd() & rnd() : iter % 4 == 1 ? rnd() | rnd() : iter % 4 == 2 ? rnd()
                              : (iter & 8 ? 0xFFFFFFFFu : (iter & 16 ? 0u : 1u << (iter % 32)));
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
    if (!pair_bricks_fit(64, 48, 32, 16) || pair_bricks_fit(48, 48, 32, 16) || pair_bricks_fit(64, 48, 32, 8)) ++bad;
    bad += check_pair_bricks(32, 16, 16);
    bad += check_pair_bricks(64, 48, 32);
    bad += check_pair_bricks(96, 32, 48);
    std::printf("bits: %d mismatches\n", bad);
    return bad ? 1 : 0;
}
