// Arguments and launchers of the fused per-level lifting kernels (fused.hip).
#pragma once

#include "common.hpp"

namespace wz {

template <class T>
struct FwdLevel {
    const T* src;       // level input box (x fastest)
    uint64_t spy, spz;  // its row / plane pitch (elements)
    Dims n;             // level box extents
    T* hi;              // coefficient volume (seven high bands)
    uint64_t hpy, hpz;
    T* lo;              // LLL destination (next level's input box)
    uint64_t lpy, lpz;
    uint32_t sz;        // z segment length (even)
    uint32_t* status;   // non-null: screen the input for non-finite samples
};

template <class T>
struct InvLevel {
    const T* hi;        // coefficient volume (high bands of this level)
    uint64_t hpy, hpz;
    const T* lo;        // reconstructed LLL box of the coarser level
    uint64_t lpy, lpz;
    Dims n;             // level box extents
    T* out;             // reconstructed level box
    uint64_t opy, opz;
    uint32_t sz;        // z segment length (multiple of 16)
};

template <class T>
void launch_fwd_level(const FwdLevel<T>& a, hipStream_t s);
template <class T>
void launch_inv_level(const InvLevel<T>& a, hipStream_t s);

}  // namespace wz
