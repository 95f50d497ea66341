// Morton bricks of a coefficient volume (cpu/src/codec.cpp:96-150): brick
// enumeration and packed base, and the quantization of one brick, shared by
// the single-field and batched pipelines.
#pragma once

#include "common.hpp"
#include "launch.hpp"

namespace wz {

struct Brick {
    uint32_t x0, y0, z0, bx, by, bz, shape;
    uint64_t base;
};

// Position (x, y, z) in brick k of its local row-major index i: shifts for
// bricks full in x and y (all but the edge ones), integer division otherwise.
WZ_D void brick_xyz(const Brick& k, uint32_t i, uint32_t& x, uint32_t& y, uint32_t& z) {
    if (k.bx == 16 && k.by == 16) {
        x = i & 15u;
        y = (i >> 4) & 15u;
        z = i >> 8;
    } else {
        const uint32_t r = i / k.bx;
        x = i - r * k.bx;
        y = r % k.by;
        z = r / k.by;
    }
}

WZ_D Brick brick_of(const BrickGeometry& g, uint32_t b) {
    Brick k;
    const uint32_t bxi = b % g.nbx, r = b / g.nbx, byi = r % g.nby, bzi = r / g.nby;
    k.x0 = bxi * 16;
    k.y0 = byi * 16;
    k.z0 = bzi * g.bz;
    k.bx = min(16u, g.d.nx - k.x0);
    k.by = min(16u, g.d.ny - k.y0);
    k.bz = min(g.bz, g.d.nz - k.z0);
    k.shape = unsigned(k.bx != 16) | (unsigned(k.by != 16) << 1) | (unsigned(k.bz != g.bz) << 2);
    k.base = uint64_t(k.z0) * g.d.nx * g.d.ny + uint64_t(k.y0) * g.d.nx * k.bz + uint64_t(k.x0) * k.by * k.bz;
    return k;
}

// Brick b of a field (one workgroup of 256 threads): coefficients quantized
// and written back dequantized, u16 magnitudes at their packed ranks (rank(k,
// i) of the brick's local row-major index i), sign bits, q-outliers appended.
template <class T, class Rank>
WZ_D void quantize_brick(T* __restrict__ coeff, const BrickGeometry& g, T q, T inv, const Rank& rank,
                         uint16_t* __restrict__ mag, uint32_t* __restrict__ sign, const SideList& qo,
                         uint32_t* __restrict__ status, uint32_t b) {
    __shared__ uint32_t bits[257];
    const Brick k = brick_of(g, b);
    const uint32_t vol = k.bx * k.by * k.bz;
    const uint64_t plane = uint64_t(g.d.nx) * g.d.ny;
    for (uint32_t w = threadIdx.x; w < 257; w += 256) bits[w] = 0;
    __syncthreads();
    const uint32_t shift = uint32_t(k.base & 31u);
    bool bad = false;
    // a burst of kBurst samples per thread: their coefficient and rank loads
    // are all issued before the first store (the stores could alias the
    // loads of the next samples, so the compiler would not hoist them)
    constexpr uint32_t kBurst = 8;
    for (uint32_t i0 = 0; i0 < vol; i0 += kBurst * 256u) {
      T cv[kBurst];
      uint32_t rv[kBurst];
      uint64_t av[kBurst];
#pragma unroll
      for (uint32_t u = 0; u < kBurst; ++u) {
        const uint32_t i = i0 + u * 256u + threadIdx.x;
        const uint32_t ic = i < vol ? i : 0u;
        uint32_t x, y, z;
        brick_xyz(k, ic, x, y, z);
        av[u] = (uint64_t(k.z0 + z) * g.d.ny + (k.y0 + y)) * g.d.nx + k.x0 + x;
        rv[u] = rank(k, ic, x, y, z);
        cv[u] = coeff[av[u]];
      }
#pragma unroll
      for (uint32_t u = 0; u < kBurst; ++u) {
        if (i0 + u * 256u + threadIdx.x >= vol) continue;
        const uint64_t at = av[u];
        const uint32_t r = rv[u];
        const T scaled = cv[u] * inv;
        // cpu/src/quantize_avx2.cpp:45-62: finite, |scaled| < 2^63, llrint, INT32 saturation.
        if (!__builtin_isfinite(scaled) || !(abs_t(scaled) < T(9223372036854775808.0))) bad = true;
        const T rounded = rint_t(scaled);
        const bool negative = rounded < T(0);
        const T am = abs_t(rounded);
        const int32_t m = am >= T(2147483648.0) ? INT32_MAX : int32_t(am);
        coeff[at] = m ? T(negative ? -m : m) * q : T(0);
        mag[k.base + r] = uint16_t(m > 65535 ? 65535 : m);
        if (negative && m) atomicOr(&bits[(r + shift) >> 5], 1u << ((r + shift) & 31u));
        append_qoutlier(qo, m > 65535, uint32_t(at), uint32_t(m));
      }
    }
    if (bad) atomicOr(status, uint32_t(kQuantDomain));
    __syncthreads();
    // Brick bits occupy global bits [base, base+vol); LDS word w maps to global
    // word base/32 + w. The first and last words may be shared with neighbours.
    const uint32_t words = (shift + vol + 31) >> 5;
    const uint64_t w0 = k.base >> 5;
    for (uint32_t w = threadIdx.x; w < words; w += 256) {
        const uint32_t v = bits[w];
        if (w == 0 || w == words - 1) {
            if (v) atomicOr(&sign[w0 + w], v);
        } else {
            sign[w0 + w] = v;
        }
    }
}


}  // namespace wz
