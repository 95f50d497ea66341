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

// Threads of a workgroup that runs lds_pass.
constexpr uint32_t kLdsThreads = 512;

// One lifting pass in place over lines of n samples in LDS (line l starts at
// l * pitch with unit stride for Ax = 0, at l with stride `pitch` otherwise).
// Lines go in groups small enough that every thread holds the (even, odd)
// pairs of its K tiles: all of a group's loads happen before its stores
// (the band split / interleave writes inside the same lines); groups touch
// disjoint lines. Consecutive lanes take consecutive lines of one tile
// column, so with an odd row pitch they hit distinct LDS banks.
template <class T, bool Inv, int Ax, int Tl>
WZ_D void lds_pass(T* A, uint32_t pitch, uint32_t n, uint32_t lines) {
    constexpr int P = Tl / 2 + 4;
    constexpr int K = (sizeof(T) == 4 ? 4 : 2) / (Tl > 8 ? 2 : 1);  // <= 128 VGPRs: two workgroups per CU
    const uint32_t t = threadIdx.x, nt = (n + Tl - 1) / Tl;
    const uint32_t per = (K * kLdsThreads) / nt;  // lines of a group (>= 1: nt <= K * threads)
    const uint32_t h = half_len(n);
    for (uint32_t l0 = 0; l0 < lines; l0 += per) {
        const uint32_t gl = min(per, lines - l0), items = gl * nt;
        const FastDiv by_gl(gl);  // w < K * threads, gl <= that: exact
        T e[K][P], o[K][P];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const uint32_t w = t + uint32_t(k) * kLdsThreads;
            if (w >= items) continue;
            const uint32_t tile = by_gl.div(w), line = l0 + (w - tile * gl);
            const uint32_t base = Ax == 0 ? line * pitch : line, stride = Ax == 0 ? 1u : pitch;
            const int32_t start = int32_t(tile) * Tl;
#pragma unroll
            for (int j = 0; j < P; ++j) {
                const int32_t q = start - 4 + 2 * j;
                e[k][j] = A[base + (Inv ? split_pos(q, n) : reflect_pos(q, n)) * stride];
                o[k][j] = A[base + (Inv ? split_pos(q + 1, n) : reflect_pos(q + 1, n)) * stride];
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const uint32_t w = t + uint32_t(k) * kLdsThreads;
            if (w >= items) continue;
            const uint32_t tile = by_gl.div(w), line = l0 + (w - tile * gl);
            const uint32_t base = Ax == 0 ? line * pitch : line, stride = Ax == 0 ? 1u : pitch;
            const uint32_t start = tile * Tl;
            if constexpr (Inv) lift_inverse<T, P, Ax != 0>(e[k], o[k]);
            else lift_forward<T, P>(e[k], o[k]);
#pragma unroll
            for (int j = 2; j < P - 2; ++j) {
                const uint32_t q = start + 2u * uint32_t(j - 2);
                if (q < n) A[base + (Inv ? q : q >> 1) * stride] = e[k][j];
                if (q + 1 < n) A[base + (Inv ? q + 1 : h + (q >> 1)) * stride] = o[k][j];
            }
        }
        __syncthreads();
    }
}

}  // namespace wz
