// CDF 9/7 lifting of P (even, odd) sample pairs held in registers, shared by
// the axis-pass kernels of the single-field and batched paths.
#pragma once

#include "common.hpp"

namespace wz {

// Forward lifting of P pairs in registers (cpu/src/transform.cpp:136-145).
template <class T, int P>
WZ_D void lift_forward(T (&e)[P], T (&o)[P]) {
    using C = Lift<T>;
#pragma unroll
    for (int j = 0; j < P - 1; ++j) o[j] = fmadd(C::a, e[j] + e[j + 1], o[j]);
#pragma unroll
    for (int j = 1; j < P - 1; ++j) e[j] = fmadd(C::b, o[j - 1] + o[j], e[j]);
#pragma unroll
    for (int j = 1; j < P - 2; ++j) o[j] = fmadd(C::g, e[j] + e[j + 1], o[j]);
#pragma unroll
    for (int j = 2; j < P - 2; ++j) e[j] = fmadd(C::d, o[j - 1] + o[j], e[j]);
#pragma unroll
    for (int j = 2; j < P - 2; ++j) {
        e[j] = e[j] * C::k;
        o[j] = o[j] * (-C::ik);
    }
}

// Inverse lifting of P pairs; Folded contracts the two tile-end sums
// (cpu/src/transform.cpp:117-135).
template <class T, int P, bool Folded>
WZ_D void lift_inverse(T (&e)[P], T (&o)[P]) {
    using C = Lift<T>;
    const T first_odd = o[0], last_odd = o[P - 1];
#pragma unroll
    for (int j = 0; j < P; ++j) {
        e[j] = e[j] * C::ik;
        o[j] = o[j] * (-C::k);
    }
#pragma unroll
    for (int j = 1; j < P; ++j) {
        T s;
        if (Folded && j == 1) s = fmadd(first_odd, -C::k, o[1]);
        else if (Folded && j == P - 1) s = fmadd(last_odd, -C::k, o[P - 2]);
        else s = o[j - 1] + o[j];
        e[j] = fmadd(-C::d, s, e[j]);
    }
#pragma unroll
    for (int j = 1; j < P - 1; ++j) o[j] = fmadd(-C::g, e[j] + e[j + 1], o[j]);
#pragma unroll
    for (int j = 2; j < P - 1; ++j) e[j] = fmadd(-C::b, o[j - 1] + o[j], e[j]);
#pragma unroll
    for (int j = 2; j < P - 2; ++j) o[j] = fmadd(-C::a, e[j] + e[j + 1], o[j]);
}

template <class T>
WZ_D bool finite_value(T v) {
    return __builtin_isfinite(v);
}

// One tile (Tile outputs) of one line: the line's n samples at src[base +
// k * stride], results to dst[dbase + k * dstride] (the band split of the
// forward, the interleave of the inverse). Generic pointers: global or LDS.
template <class T, bool Inverse, int Axis, int Tile>
WZ_D void lift_tile(const T* src, uint64_t base, uint64_t stride, T* dst, uint64_t dbase, uint64_t dstride,
                    uint32_t n, uint32_t t, uint32_t* status) {
    constexpr int P = Tile / 2 + 4;
    const int32_t start = int32_t(t) * Tile;
    T e[P], o[P];
#pragma unroll
    for (int j = 0; j < P; ++j) {
        const int32_t p = start - 4 + 2 * j;
        const uint32_t ia = Inverse ? split_pos(p, n) : reflect_pos(p, n);
        const uint32_t ib = Inverse ? split_pos(p + 1, n) : reflect_pos(p + 1, n);
        e[j] = src[base + uint64_t(ia) * stride];
        o[j] = src[base + uint64_t(ib) * stride];
    }
    if (status) {  // input screen on the first forward pass (cpu/src/codec.cpp:183)
        bool ok = true;
#pragma unroll
        for (int j = 0; j < P; ++j) ok = ok && finite_value(e[j]) && finite_value(o[j]);
        if (!ok) atomicOr(status, uint32_t(kNonFiniteInput));
    }
    if constexpr (Inverse) lift_inverse<T, P, Axis != 0>(e, o);
    else lift_forward<T, P>(e, o);
    const uint32_t h = half_len(n);
#pragma unroll
    for (int j = 2; j < P - 2; ++j) {
        const uint32_t p = uint32_t(start) + 2u * uint32_t(j - 2);
        if (p < n) dst[dbase + uint64_t(Inverse ? p : p >> 1) * dstride] = e[j];
        if (p + 1 < n) dst[dbase + uint64_t(Inverse ? p + 1 : h + (p >> 1)) * dstride] = o[j];
    }
}

}  // namespace wz
