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

}  // namespace wz
