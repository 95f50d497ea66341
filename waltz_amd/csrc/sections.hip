// Side channels (native repairs, quantization outliers) and archive assembly.
//
// The canonical archive stores both side channels in increasing index order
// (cpu/src/codec.cpp:225-257). Two producers, two atomic-free orderings:
//  * repairs come from a row-major pass: each workgroup owns one contiguous
//    tile of 4096 indices, compacts its repairs in order into the tile's own
//    segment (block scan, no global atomics), and an exclusive scan over the
//    tile counts gives every entry its final position;
//  * q-outliers come from Morton bricks (not row-major): each sets its bit in a
//    row-major flag bitmap and is appended to an unsorted list; since indices
//    are unique, an entry's sorted position is the number of flagged indices
//    below it (per-1024-bit segment popcounts + one exclusive scan + in-segment
//    popcounts). O(n), no sort, same bytes as a comparison sort.
#include "sections.hpp"

namespace wz {

template <class T>
__global__ void __launch_bounds__(256) repair_kernel(const T* __restrict__ original, const T* __restrict__ rec,
                                                     uint32_t n, double bound, SegList r,
                                                     uint32_t* __restrict__ status) {
    repair_tile(original, rec, n, bound, r, blockIdx.x, status);
}

template <class T>
void detect_repairs(const T* original, const T* reconstructed, uint32_t n, double bound, SegList r,
                    uint32_t* status, hipStream_t s) {
    const uint32_t tiles = (n + kRepairTile - 1) / kRepairTile;
    hipLaunchKernelGGL((repair_kernel<T>), dim3(tiles), dim3(256), 0, s, original, reconstructed, n, bound, r,
                       status);
}

// Word prefixes of both flag bitmaps in one launch: workgroups [0, blocks)
// take the repair bitmap, [blocks, 2 blocks) the q-outlier bitmap (skipped
// when there are no q-outliers).
__global__ void __launch_bounds__(256) word_prefix2(const uint32_t* __restrict__ rbm, uint32_t* __restrict__ rwpre,
                                                    uint32_t* __restrict__ rbsum, const uint32_t* __restrict__ qbm,
                                                    uint32_t* __restrict__ qwpre, uint32_t* __restrict__ qbsum,
                                                    const uint32_t* __restrict__ qcount, uint32_t words,
                                                    uint32_t blocks) {
    if (blockIdx.x < blocks) {
        word_prefix_block(rbm, words, rwpre, rbsum, blockIdx.x);
    } else if (*qcount != 0) {
        word_prefix_block(qbm, words, qwpre, qbsum, blockIdx.x - blocks);
    }
}

template <class T>
__global__ void __launch_bounds__(1024) scan_layout_kernel(uint8_t* __restrict__ archive, uint64_t capacity,
                                                           uint32_t n, uint32_t* rbsum, uint32_t* qbsum,
                                                           uint32_t blocks, uint32_t* np_dev, SideList qo,
                                                           uint8_t flags, uint8_t lx, uint8_t lz, uint8_t bz,
                                                           double quantum, EncodeResult* __restrict__ res) {
    scan_layout<T>(archive, capacity, n, rbsum, qbsum, blocks, np_dev, qo, flags, lx, lz, bz, quantum, res);
}

// Both side channels to their index-order slots in one launch: item
// g < segments * kPlaceSlices places repairs of segment g / kPlaceSlices,
// the remaining kPlaceQBlocks place q-outliers. Items map to workgroups
// XCD-aware: the dispatcher deals workgroups round-robin to the 8 XCDs, and
// XCD k takes the k-th contiguous range of items instead, so the slices of
// one segment and the neighbouring segments whose entries share output rows
// (segments follow the finest level's x-fastest tile order) write those
// lines, and read those bitmap words, through one L2. Grid = 8 * ceil(items
// / 8); spare workgroups return at once.
template <class T>
__global__ void __launch_bounds__(256) place_kernel(uint8_t* __restrict__ archive, SegList rp, Ranks rr, SideList qo,
                                                    Ranks qr, const EncodeResult* __restrict__ res) {
    if (res->status) return;
    const uint32_t rblocks = rp.segments * kPlaceSlices, per = gridDim.x / 8;
    const uint32_t g = (blockIdx.x & 7u) * per + (blockIdx.x >> 3);
    if (g >= rblocks + kPlaceQBlocks) return;
    if (g < rblocks) place_repairs<T>(archive, rp, rr, res, g / kPlaceSlices, g % kPlaceSlices);
    else place_qoutliers(archive, qo, qr, uint64_t(sizeof(T)), res, g - rblocks);
}

// Scratch layout: [repair wpre | repair bsum | q wpre | q bsum], each part a
// multiple of 64 words so that the 16-byte vector accesses stay aligned.
namespace {
uint64_t pad64(uint64_t w) { return (w + 63) & ~uint64_t(63); }
}  // namespace
uint64_t finish_scratch_words(uint32_t n) {
    const uint64_t words = (uint64_t(n) + 31) / 32, blocks = (words + kPrefixWords - 1) / kPrefixWords;
    return 2 * (pad64(words) + pad64(blocks));
}

template <class T>
void finish_archive(uint8_t* archive, uint64_t capacity, uint32_t n, SegList repairs, SideList qout, uint32_t* scratch,
                    uint8_t flags, uint8_t lx, uint8_t lz, uint8_t bz, double quantum, EncodeResult* result,
                    hipStream_t s) {
    const uint32_t words = (n + 31) / 32, blocks = (words + kPrefixWords - 1) / kPrefixWords;
    const uint64_t half = pad64(words) + pad64(blocks);
    const Ranks rr{repairs.bitmap, scratch, scratch + pad64(words)};
    const Ranks qr{qout.bitmap, scratch + half, scratch + half + pad64(words)};
    uint32_t* np_total = &result->pad;  // repair total from the bitmap scan
    hipLaunchKernelGGL(word_prefix2, dim3(2 * blocks), dim3(256), 0, s, repairs.bitmap, const_cast<uint32_t*>(rr.wpre),
                       const_cast<uint32_t*>(rr.bsum), qout.bitmap, const_cast<uint32_t*>(qr.wpre),
                       const_cast<uint32_t*>(qr.bsum), static_cast<const uint32_t*>(qout.count), words, blocks);
    hipLaunchKernelGGL((scan_layout_kernel<T>), dim3(1), dim3(1024), 0, s, archive, capacity, n,
                       const_cast<uint32_t*>(rr.bsum), const_cast<uint32_t*>(qr.bsum), blocks, np_total, qout, flags,
                       lx, lz, bz, quantum, result);
    const uint32_t items = repairs.segments * kPlaceSlices + kPlaceQBlocks;
    hipLaunchKernelGGL((place_kernel<T>), dim3(8 * ((items + 7) / 8)), dim3(256), 0, s,
                       archive, repairs, rr, qout, qr, result);
}

// ------------------------------------------------------------------ decode patches
// q-outlier record j: the coefficient's magnitude replaced (sign kept).
template <class T>
WZ_D void qpatch_one(T* __restrict__ c, const uint32_t* __restrict__ idx, const uint32_t* __restrict__ val,
                     uint32_t j, uint32_t n, T q, uint32_t* __restrict__ status) {
    const uint32_t i = idx[j], v = val[j];
    if (i >= n || v <= 65535u || v > 0x7FFFFFFFu) {
        atomicOr(status, uint32_t(kCorrupt));
        return;
    }
    if (j && idx[j - 1] >= i) atomicOr(status, uint32_t(kUnordered));  // duplicates are checked apart
    const T x = T(v) * q;  // cpu/src/codec.cpp:299-302
    c[i] = c[i] < T(0) ? -x : x;
}

template <class T>
__global__ void qpatch_kernel(T* __restrict__ c, const uint32_t* __restrict__ idx, const uint32_t* __restrict__ val,
                              uint32_t count, uint32_t n, T q, uint32_t* __restrict__ status) {
    const uint32_t j = blockIdx.x * 256u + threadIdx.x;
    if (j < count) qpatch_one(c, idx, val, j, n, q, status);
}

// Repair record j: the sample's correction subtracted, with the record checks.
template <class T>
WZ_D void rpatch_one(T* __restrict__ out, const uint32_t* __restrict__ idx, const T* __restrict__ err, uint32_t j,
                     uint32_t n, uint32_t* __restrict__ status) {
    const uint32_t i = idx[j];
    const T e = err[j];
    if (i >= n || !__builtin_isfinite(e)) {
        atomicOr(status, uint32_t(kCorrupt));
        return;
    }
    if (j && idx[j - 1] >= i) atomicOr(status, uint32_t(kUnordered));
    const T v = out[i] - e;  // cpu/src/codec.cpp:330
    out[i] = v;
    if (!__builtin_isfinite(v)) atomicOr(status, uint32_t(kNonFiniteOutput));
}

template <class T>
__global__ void rpatch_kernel(T* __restrict__ out, const uint32_t* __restrict__ idx, const T* __restrict__ err,
                              uint32_t count, uint32_t n, uint32_t* __restrict__ status) {
    const uint32_t j = blockIdx.x * 256u + threadIdx.x;
    if (j < count) rpatch_one(out, idx, err, j, n, status);
}

// Repair record j: the checks of rpatch_kernel, and the record that opens
// its (row, 64-sample x tile) key writes its position to the table.
template <class T>
WZ_D void repair_key_one(const uint32_t* __restrict__ idx, const T* __restrict__ err, uint32_t j, uint32_t n,
                         uint32_t nx, uint32_t txn, uint32_t* __restrict__ table, uint32_t* __restrict__ status) {
    const uint32_t i = idx[j];
    if (i >= n || !__builtin_isfinite(err[j])) {
        atomicOr(status, uint32_t(kCorrupt));
        return;
    }
    auto key = [&](uint32_t v) { return (v / nx) * txn + (v % nx) / 64u; };
    bool opens = true;
    if (j) {
        const uint32_t p = idx[j - 1];
        if (p >= i) atomicOr(status, uint32_t(kUnordered));
        opens = p >= n || key(p) != key(i);
    }
    if (opens) table[key(i)] = j;
}

template <class T>
__global__ void repair_keys_kernel(const uint32_t* __restrict__ idx, const T* __restrict__ err, uint32_t count,
                                   uint32_t n, uint32_t nx, uint32_t txn, uint32_t* __restrict__ table,
                                   uint32_t* __restrict__ status) {
    const uint32_t j = blockIdx.x * 256u + threadIdx.x;
    if (j < count) repair_key_one(idx, err, j, n, nx, txn, table, status);
}

// ------------------------------------------------------------------ planned decode
// The header check of a decode planned on the device (one thread).
__global__ void decode_plan_kernel(const uint8_t* __restrict__ archive, uint64_t bytes, PlanExpect e,
                                   DecodeRecord* __restrict__ rec) {
    if (threadIdx.x == 0) plan_decode(archive, bytes, e, rec);
}

// The list kernels of a planned decode: a fixed grid walks the record's counts.
constexpr uint32_t kPlannedBlocks = 512;

template <class T>
__global__ void __launch_bounds__(256) qpatch_planned_kernel(T* __restrict__ c, const uint8_t* __restrict__ archive,
                                                             const DecodeRecord* __restrict__ rec, uint32_t n,
                                                             uint32_t* __restrict__ status) {
    if (!rec->ok) return;
    const uint32_t count = rec->nq;
    const auto* idx = reinterpret_cast<const uint32_t*>(archive + rec->qi);
    const auto* val = reinterpret_cast<const uint32_t*>(archive + rec->qv);
    const T q = T(rec->quantum);
    for (uint32_t j = blockIdx.x * 256u + threadIdx.x; j < count; j += gridDim.x * 256u)
        qpatch_one(c, idx, val, j, n, q, status);
}

template <class T>
__global__ void __launch_bounds__(256) rpatch_planned_kernel(T* __restrict__ out, const uint8_t* __restrict__ archive,
                                                             const DecodeRecord* __restrict__ rec, uint32_t n,
                                                             uint32_t* __restrict__ status) {
    if (!rec->ok) return;
    const uint32_t count = rec->np;
    const auto* idx = reinterpret_cast<const uint32_t*>(archive + rec->pi);
    const auto* err = reinterpret_cast<const T*>(archive + rec->pv);
    for (uint32_t j = blockIdx.x * 256u + threadIdx.x; j < count; j += gridDim.x * 256u)
        rpatch_one(out, idx, err, j, n, status);
}

template <class T>
__global__ void __launch_bounds__(256) repair_keys_planned_kernel(const uint8_t* __restrict__ archive,
                                                                  const DecodeRecord* __restrict__ rec, uint32_t n,
                                                                  uint32_t nx, uint32_t txn,
                                                                  uint32_t* __restrict__ table,
                                                                  uint32_t* __restrict__ status) {
    if (!rec->ok) return;
    const uint32_t count = rec->np;
    const auto* idx = reinterpret_cast<const uint32_t*>(archive + rec->pi);
    const auto* err = reinterpret_cast<const T*>(archive + rec->pv);
    for (uint32_t j = blockIdx.x * 256u + threadIdx.x; j < count; j += gridDim.x * 256u)
        repair_key_one(idx, err, j, n, nx, txn, table, status);
}

// Duplicate detection of an unordered side list (cpu/src/codec.cpp:260-275).
__global__ void __launch_bounds__(256) unique_kernel(const uint32_t* __restrict__ idx, uint32_t count, uint32_t n,
                                                     uint32_t* __restrict__ bitmap, uint32_t* __restrict__ status) {
    const uint32_t j = blockIdx.x * 256u + threadIdx.x;
    if (j >= count) return;
    const uint32_t i = idx[j];
    if (i >= n) {
        atomicOr(status, uint32_t(kCorrupt));
        return;
    }
    const uint32_t bit = 1u << (i & 31u);
    if (atomicOr(&bitmap[i >> 5], bit) & bit) atomicOr(status, uint32_t(kCorrupt));
}

template <class T>
__global__ void __launch_bounds__(256) finite_kernel(const T* __restrict__ v, uint32_t n, uint32_t* __restrict__ status) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    const bool bad = i < n && !__builtin_isfinite(v[i]);
    if (__ballot(bad) && sec::lane_id() == 0) atomicOr(status, uint32_t(kNonFiniteOutput));
}

template <class T>
void apply_qoutliers(T* coeff, const uint8_t* idx, const uint8_t* val, uint32_t count, uint32_t n, double quantum,
                     uint32_t* status, hipStream_t s) {
    if (!count) return;
    hipLaunchKernelGGL((qpatch_kernel<T>), dim3((count + 255) / 256), dim3(256), 0, s, coeff,
                       reinterpret_cast<const uint32_t*>(idx), reinterpret_cast<const uint32_t*>(val), count, n,
                       T(quantum), status);
}

template <class T>
void apply_repairs(T* out, const uint8_t* idx, const uint8_t* err, uint32_t count, uint32_t n, uint32_t* status,
                   hipStream_t s) {
    if (!count) return;
    hipLaunchKernelGGL((rpatch_kernel<T>), dim3((count + 255) / 256), dim3(256), 0, s, out,
                       reinterpret_cast<const uint32_t*>(idx), reinterpret_cast<const T*>(err), count, n, status);
}

template <class T>
void apply_repairs_planned(T* out, const uint8_t* archive, const DecodeRecord* rec, uint32_t n, uint32_t* status,
                           hipStream_t s) {
    hipLaunchKernelGGL((rpatch_planned_kernel<T>), dim3(kPlannedBlocks), dim3(256), 0, s, out, archive, rec, n,
                       status);
}

template <class T>
void repair_keys(const uint32_t* idx, const T* err, uint32_t count, Dims d, uint32_t* table, uint32_t* status,
                 hipStream_t s) {
    if (!count) return;
    const uint32_t n = d.nx * d.ny * d.nz;
    hipLaunchKernelGGL((repair_keys_kernel<T>), dim3((count + 255) / 256), dim3(256), 0, s, idx, err, count, n, d.nx,
                       (d.nx + 63) / 64, table, status);
}

void decode_plan(const uint8_t* archive, uint64_t archive_bytes, const PlanExpect& e, DecodeRecord* rec,
                 hipStream_t s) {
    hipLaunchKernelGGL(decode_plan_kernel, dim3(1), dim3(64), 0, s, archive, archive_bytes, e, rec);
}

template <class T>
void apply_qoutliers_planned(T* coeff, const uint8_t* archive, const DecodeRecord* rec, uint32_t n, uint32_t* status,
                             hipStream_t s) {
    hipLaunchKernelGGL((qpatch_planned_kernel<T>), dim3(kPlannedBlocks), dim3(256), 0, s, coeff, archive, rec, n,
                       status);
}

template <class T>
void repair_keys_planned(const uint8_t* archive, const DecodeRecord* rec, Dims d, uint32_t* table, uint32_t* status,
                         hipStream_t s) {
    hipLaunchKernelGGL((repair_keys_planned_kernel<T>), dim3(kPlannedBlocks), dim3(256), 0, s, archive, rec,
                       d.nx * d.ny * d.nz, d.nx, (d.nx + 63) / 64, table, status);
}

uint64_t patch_table_words(Dims full, uint32_t np, unsigned width) {
    const uint64_t n = uint64_t(full.nx) * full.ny * full.nz;
    const uint64_t words = uint64_t(full.ny) * full.nz * ((full.nx + 63) / 64);
    // the table sits in the decoder's magnitude scratch (2n bytes); the
    // kernel's 32-bit buffer windows hold the table and both list sections
    const uint64_t limit = uint64_t(1) << 31;
    if (!np || 4 * words > 2 * n || 4 * words > limit || uint64_t(np) * 4 > limit || uint64_t(np) * width > limit)
        return 0;
    return words;
}

void check_unique(const uint8_t* idx, uint32_t count, uint32_t n, uint32_t* bitmap, uint32_t* status, hipStream_t s) {
    if (!count) return;
    zero_words(bitmap, (uint64_t(n) + 31) / 32, s);
    hipLaunchKernelGGL(unique_kernel, dim3((count + 255) / 256), dim3(256), 0, s,
                       reinterpret_cast<const uint32_t*>(idx), count, n, bitmap, status);
}

template <class T>
void check_finite(const T* out, uint32_t n, uint32_t* status, hipStream_t s) {
    hipLaunchKernelGGL((finite_kernel<T>), dim3((n + 255) / 256), dim3(256), 0, s, out, n, status);
}

#define WZ_INST(T)                                                                                                \
    template void detect_repairs<T>(const T*, const T*, uint32_t, double, SegList, uint32_t*, hipStream_t);         \
    template void finish_archive<T>(uint8_t*, uint64_t, uint32_t, SegList, SideList, uint32_t*, uint8_t, uint8_t,  \
                                    uint8_t, uint8_t, double, EncodeResult*, hipStream_t);                         \
    template void apply_qoutliers<T>(T*, const uint8_t*, const uint8_t*, uint32_t, uint32_t, double, uint32_t*,    \
                                     hipStream_t);                                                                \
    template void apply_repairs<T>(T*, const uint8_t*, const uint8_t*, uint32_t, uint32_t, uint32_t*, hipStream_t); \
    template void repair_keys<T>(const uint32_t*, const T*, uint32_t, Dims, uint32_t*, uint32_t*, hipStream_t);     \
    template void apply_qoutliers_planned<T>(T*, const uint8_t*, const DecodeRecord*, uint32_t, uint32_t*,         \
                                             hipStream_t);                                                        \
    template void repair_keys_planned<T>(const uint8_t*, const DecodeRecord*, Dims, uint32_t*, uint32_t*,           \
                                         hipStream_t);                                                            \
    template void apply_repairs_planned<T>(T*, const uint8_t*, const DecodeRecord*, uint32_t, uint32_t*,           \
                                           hipStream_t);                                                          \
    template void check_finite<T>(const T*, uint32_t, uint32_t*, hipStream_t);
WZ_INST(float)
WZ_INST(double)
#undef WZ_INST

}  // namespace wz
