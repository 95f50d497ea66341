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

// Whether `rows` rows of pitch `pitch` elements of `width` bytes stay below
// 4 GiB less 1 MiB (the level kernels' 32-bit byte offsets within a plane,
// with room above the plane for the offset that masks a lane off).
inline bool fused_plane_fits(uint64_t pitch, uint64_t rows, uint64_t width) {
    return pitch * rows * width <= (uint64_t(1) << 32) - (uint64_t(1) << 20);
}

enum InvEmit : int {
    kEmitStore = 0,        // write the reconstruction
    kEmitStoreScreen = 1,  // write it and flag non-finite samples
    kEmitRepair = 2,       // no write: native repair records against `orig`
};

template <class T>
struct InvLevel {
    const T* hi;        // coefficient volume (high bands of this level)
    uint32_t hpy, hpz;  // pitches in elements (planes below 4 GiB)
    const T* lo;        // reconstructed LLL box of the coarser level
    uint32_t lpy, lpz;
    Dims n;             // level box extents
    T* out;             // reconstructed level box
    uint32_t opy, opz;  // (also the pitches of `orig`)
    uint32_t sz;        // z segment length (multiple of 16)
    int emit;           // InvEmit
    uint32_t* status;   // screen / strict-bound failures
    const T* orig;      // kEmitRepair: the original field
    double bound;
    float bound_lo;     // float repair emitter: the largest float <= bound
    T q, inv;           // q > 0: high-band loads are quantized and dequantized (repair emitter, store)
    SegList rep;        // kEmitRepair: one segment per workgroup, cap = 64*32*sz
    // kEmitStoreScreen, np > 0: repairs applied as the samples are stored
    // (strictly increasing indices; key table from repair_keys, `keys` words)
    const uint32_t* pidx;
    const T* perr;
    const uint32_t* ptable;
    uint32_t np, keys;
    const DecodeRecord* plan;  // non-null: np and the list's places from this record of `archive`
    const uint8_t* archive;
};

// Repair segments needed by a level-0 inverse with z segment sz.
inline uint64_t inv_segments(Dims n, uint32_t sz) {
    return uint64_t((n.nx + 63) / 64) * ((n.ny + 31) / 32) * ((n.nz + sz - 1) / sz);
}

template <class T>
void launch_fwd_level(const FwdLevel<T>& a, hipStream_t s);
template <class T>
void launch_inv_level(const InvLevel<T>& a, hipStream_t s);
// z segment of the finest level's repair emitter for box n: the tuned segment,
// doubled while the tiles outnumber the workgroups the device holds at once
// (a multiple of 16 either way)
template <class T>
uint32_t repair_z_segment(Dims n, uint32_t tuned);

}  // namespace wz
