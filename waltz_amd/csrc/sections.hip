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
#include "common.hpp"
#include "launch.hpp"

namespace wz {

namespace {
WZ_D uint32_t lane_id() { return threadIdx.x & 63u; }

WZ_D uint32_t wave_inclusive(uint32_t v) {
#pragma unroll
    for (uint32_t d = 1; d < 64; d <<= 1) {
        const uint32_t t = __shfl_up(v, d, 64);
        if (lane_id() >= d) v += t;
    }
    return v;
}

// Exclusive scan over a 256-thread block; `sh` >= 4 words.
WZ_D uint32_t block_exclusive256(uint32_t v, uint32_t* sh, uint32_t& total) {
    const uint32_t inc = wave_inclusive(v);
    const uint32_t w = threadIdx.x >> 6;
    if (lane_id() == 63) sh[w] = inc;
    __syncthreads();
    uint32_t before = 0, sum = 0;
#pragma unroll
    for (uint32_t i = 0; i < 4; ++i) {
        before += i < w ? sh[i] : 0u;
        sum += sh[i];
    }
    __syncthreads();
    total = sum;
    return before + inc - v;
}

}  // namespace

// Native repairs (cpu/src/repair_avx2.cpp:58-74) over row-major tiles of
// kRepairTile indices (one segment of the SegList each). Round k of tile t
// covers indices t*4096 + 256k + [0,256) (one coalesced load per wave per
// round); each wave owns the two flag-bitmap words of its 64 indices, and the
// tile's entries are compacted in index order through ballot ranks and a
// 64-entry scan in LDS.
template <class T>
__global__ void __launch_bounds__(256) repair_kernel(const T* __restrict__ original, const T* __restrict__ rec,
                                                     uint32_t n, double bound, SegList r,
                                                     uint32_t* __restrict__ status) {
    constexpr uint32_t rounds = kRepairTile / 256;
    __shared__ uint32_t cnt[rounds * 4];
    const uint64_t tile0 = uint64_t(blockIdx.x) * kRepairTile;
    const uint32_t lane = lane_id(), wv = threadIdx.x >> 6;
    const unsigned long long below = (1ull << lane) - 1ull;
    T err[rounds];
    unsigned long long ballots[rounds];
    bool ok = true;
    const T nb = T(bound);
#pragma unroll
    for (uint32_t k = 0; k < rounds; ++k) {
        const uint64_t i = tile0 + 256u * k + threadIdx.x;
        bool patch = false;
        err[k] = T(0);
        if (i < n) patch = repair_decision<T>(rec[i], original[i], nb, bound, &err[k], &ok);
        ballots[k] = __ballot(patch);
        const uint64_t i0 = i - lane;  // multiple of 64
        if (lane == 0 && i0 < n) {
            r.bitmap[i0 >> 5] = uint32_t(ballots[k]);
            if (i0 + 32 < n) r.bitmap[(i0 >> 5) + 1] = uint32_t(ballots[k] >> 32);
        }
        if (lane == 0) cnt[4 * k + wv] = uint32_t(__popcll(ballots[k]));
    }
    __syncthreads();
    if (wv == 0) {  // exclusive scan of the 64 (round, wave) counts, in index order
        const uint32_t v = cnt[lane];
        const uint32_t inc = wave_inclusive(v);
        cnt[lane] = inc - v;
        if (lane == 63) r.count[blockIdx.x] = inc;
    }
    __syncthreads();
    uint32_t* idx = r.index + uint64_t(blockIdx.x) * r.cap;
    T* val = static_cast<T*>(r.payload) + uint64_t(blockIdx.x) * r.cap;
#pragma unroll
    for (uint32_t k = 0; k < rounds; ++k)
        if ((ballots[k] >> lane) & 1ull) {
            const uint32_t at = cnt[4 * k + wv] + uint32_t(__popcll(ballots[k] & below));
            idx[at] = uint32_t(tile0 + 256u * k + threadIdx.x);
            val[at] = err[k];
        }
    if (__ballot(!ok) && lane == 0) atomicOr(status, uint32_t(kRepairBound));
}

template <class T>
void detect_repairs(const T* original, const T* reconstructed, uint32_t n, double bound, SegList r,
                    uint32_t* status, hipStream_t s) {
    const uint32_t tiles = (n + kRepairTile - 1) / kRepairTile;
    hipLaunchKernelGGL((repair_kernel<T>), dim3(tiles), dim3(256), 0, s, original, reconstructed, n, bound, r,
                       status);
}

// ------------------------------------------------------------------ ordering
// Index-order rank of a flagged sample: every flag bitmap word gets the count
// of flags before it inside its block of kPrefixWords words (wpre), and every
// block the count before it (bsum, scanned); then rank = bsum[block] +
// wpre[word] + popcount of the lower bits of the word.
constexpr uint32_t kPrefixWords = 8192;  // 256 threads x 32 words

WZ_D void word_prefix_block(const uint32_t* __restrict__ bm, uint32_t words, uint32_t* __restrict__ wpre,
                             uint32_t* __restrict__ bsum, uint32_t block) {
    __shared__ uint32_t sh[4];
    const uint64_t w0 = uint64_t(block) * kPrefixWords + 32u * threadIdx.x;
    uint32_t c[32];
    if (w0 + 32 <= words) {
        const uint4* v = reinterpret_cast<const uint4*>(bm + w0);
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const uint4 x = v[q];
            c[4 * q] = __builtin_popcount(x.x);
            c[4 * q + 1] = __builtin_popcount(x.y);
            c[4 * q + 2] = __builtin_popcount(x.z);
            c[4 * q + 3] = __builtin_popcount(x.w);
        }
    } else {
#pragma unroll
        for (int i = 0; i < 32; ++i) c[i] = w0 + i < words ? uint32_t(__builtin_popcount(bm[w0 + i])) : 0u;
    }
    uint32_t sum = 0;
#pragma unroll
    for (int i = 0; i < 32; ++i) sum += c[i];
    uint32_t total;
    uint32_t run = block_exclusive256(sum, sh, total);
    if (w0 + 32 <= words) {
        uint4* o = reinterpret_cast<uint4*>(wpre + w0);
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            uint4 x;
            x.x = run;
            run += c[4 * q];
            x.y = run;
            run += c[4 * q + 1];
            x.z = run;
            run += c[4 * q + 2];
            x.w = run;
            run += c[4 * q + 3];
            o[q] = x;
        }
    } else {
#pragma unroll
        for (int i = 0; i < 32; ++i) {
            if (w0 + i < words) wpre[w0 + i] = run;
            run += c[i];
        }
    }
    if (threadIdx.x == 0) bsum[block] = total;
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

// In-place exclusive scan of a[0..len) by one 1024-thread workgroup; the sum
// goes to *total. Skipped (a untouched, *total untouched) when *gate == 0.
WZ_D void exclusive_scan_block(uint32_t* a, uint32_t len, uint32_t* total) {
    __shared__ uint32_t wsum[16];
    __shared__ uint32_t carry_sh;
    if (threadIdx.x == 0) carry_sh = 0;
    __syncthreads();
    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
    for (uint32_t base = 0; base < len; base += 1024) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t v = i < len ? a[i] : 0u;
        const uint32_t inc = wave_inclusive(v);
        if (lane == 63) wsum[wv] = inc;
        __syncthreads();
        uint32_t before = carry_sh, all = 0;
        for (uint32_t k = 0; k < 16; ++k) {
            before += k < wv ? wsum[k] : 0u;
            all += wsum[k];
        }
        if (i < len) a[i] = before + inc - v;
        __syncthreads();
        if (threadIdx.x == 0) carry_sh += all;
        __syncthreads();
    }
    if (threadIdx.x == 0 && total) *total = carry_sh;
    __syncthreads();
}

// One 1024-thread workgroup: scans both block sums (the repair total lands in
// *np_dev), then thread 0 lays out the archive sections and writes the header.
template <class T>
__global__ void __launch_bounds__(1024) scan_layout_kernel(uint8_t* __restrict__ archive, uint64_t capacity,
                                                           uint32_t n, uint32_t* rbsum, uint32_t* qbsum,
                                                           uint32_t blocks, uint32_t* np_dev, SideList qo,
                                                           uint8_t flags, uint8_t lx, uint8_t lz, uint8_t bz,
                                                           double quantum, EncodeResult* __restrict__ res) {
    exclusive_scan_block(rbsum, blocks, np_dev);
    if (*qo.count != 0) exclusive_scan_block(qbsum, blocks, nullptr);
    if (threadIdx.x != 0) return;
    const uint32_t np = *np_dev, nq = *qo.count;
    uint32_t st = res->status;
    if (nq > qo.capacity || (n > (1u << 20) && nq >= (1u << 20)) || (n > (1u << 22) && np >= (1u << 22)))
        st |= kSideCapacity;
    const uint64_t wb = res->wzp_bytes;
    const Layout L = archive_layout(wb, np, uint64_t(np) * sizeof(T), nq);
    if (L.end > capacity) st |= kWzpCapacity;
    res->np = np;
    res->nq = nq;
    res->end = L.end;
    res->status = st;
    if (st) return;
    uint8_t h[32];
    h[0] = flags;
    h[1] = lx;
    h[2] = lz;
    h[3] = bz;
    const uint32_t pb = np * uint32_t(sizeof(T));
    const uint64_t qb = __double_as_longlong(quantum);
    for (int k = 0; k < 4; ++k) h[4 + k] = uint8_t(pb >> (8 * k));
    for (int k = 0; k < 8; ++k) h[8 + k] = uint8_t(qb >> (8 * k));
    for (int k = 0; k < 8; ++k) h[16 + k] = uint8_t(wb >> (8 * k));
    for (int k = 0; k < 4; ++k) h[24 + k] = uint8_t(np >> (8 * k));
    for (int k = 0; k < 4; ++k) h[28 + k] = uint8_t(nq >> (8 * k));
    for (int k = 0; k < 32; ++k) archive[k] = h[k];
    // zero padding after each section (cpu/src/codec.cpp:247-254)
    for (uint64_t p = 32 + wb; p < L.pi; ++p) archive[p] = 0;
    for (uint64_t p = L.pi + 4ull * np; p < L.pv; ++p) archive[p] = 0;
    for (uint64_t p = L.pv + uint64_t(np) * sizeof(T); p < L.qi; ++p) archive[p] = 0;
    for (uint64_t p = L.qi + 4ull * nq; p < L.qv; ++p) archive[p] = 0;
    for (uint64_t p = L.qv + 4ull * nq; p < L.end; ++p) archive[p] = 0;
}

struct Ranks {
    const uint32_t* bm;
    const uint32_t* wpre;
    const uint32_t* bsum;  // exclusive, per kPrefixWords block
    WZ_D uint32_t of(uint32_t idx) const {
        const uint32_t w = idx >> 5;
        return bsum[w / kPrefixWords] + wpre[w] + uint32_t(__builtin_popcount(bm[w] & ((1u << (idx & 31u)) - 1u)));
    }
};

// Repairs: segment k's entries go to their index-order rank (bitmap popcounts);
// gridDim.y workgroups share a segment so that enough index -> rank -> store
// chains are in flight.
constexpr uint32_t kPlaceSlices = 8;
template <class T>
WZ_D void place_repairs(uint8_t* __restrict__ archive, SegList rp, Ranks rank, const EncodeResult* __restrict__ res,
                        uint32_t k, uint32_t slice) {
    const uint32_t c = rp.count[k];
    if (!c) return;
    const uint32_t np = res->np;
    const Layout L = archive_layout(res->wzp_bytes, np, uint64_t(np) * sizeof(T), res->nq);
    const uint32_t* idx = rp.index + uint64_t(k) * rp.cap;
    const T* val = static_cast<const T*>(rp.payload) + uint64_t(k) * rp.cap;
    uint32_t* oi = reinterpret_cast<uint32_t*>(archive + L.pi);
    T* ov = reinterpret_cast<T*>(archive + L.pv);
    for (uint32_t j = slice * 256u + threadIdx.x; j < c; j += 256u * kPlaceSlices) {
        const uint32_t i = idx[j];
        const uint32_t pos = rank.of(i);
        oi[pos] = i;
        ov[pos] = val[j];
    }
}

constexpr uint32_t kPlaceQBlocks = 256;
WZ_D void place_qoutliers(uint8_t* __restrict__ archive, SideList qo, Ranks rank, uint64_t repair_bytes_per,
                          const EncodeResult* __restrict__ res, uint32_t block) {
    const uint32_t np = res->np, nq = res->nq;
    if (!nq) return;
    const Layout L = archive_layout(res->wzp_bytes, np, uint64_t(np) * repair_bytes_per, nq);
    const uint32_t stride = kPlaceQBlocks * 256u;
    for (uint32_t j = block * 256u + threadIdx.x; j < nq; j += stride) {
        const uint32_t idx = qo.index[j];
        const uint32_t pos = rank.of(idx);
        reinterpret_cast<uint32_t*>(archive + L.qi)[pos] = idx;
        reinterpret_cast<uint32_t*>(archive + L.qv)[pos] = static_cast<const uint32_t*>(qo.payload)[j];
    }
}

// Both side channels to their index-order slots in one launch: workgroup
// g < segments * kPlaceSlices places repairs of segment g / kPlaceSlices,
// the remaining kPlaceQBlocks place q-outliers.
template <class T>
__global__ void __launch_bounds__(256) place_kernel(uint8_t* __restrict__ archive, SegList rp, Ranks rr, SideList qo,
                                                    Ranks qr, const EncodeResult* __restrict__ res) {
    if (res->status) return;
    const uint32_t g = blockIdx.x, rblocks = rp.segments * kPlaceSlices;
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
    hipLaunchKernelGGL((place_kernel<T>), dim3(repairs.segments * kPlaceSlices + kPlaceQBlocks), dim3(256), 0, s,
                       archive, repairs, rr, qout, qr, result);
}

// ------------------------------------------------------------------ decode patches
template <class T>
__global__ void qpatch_kernel(T* __restrict__ c, const uint32_t* __restrict__ idx, const uint32_t* __restrict__ val,
                              uint32_t count, uint32_t n, T q, uint32_t* __restrict__ status) {
    const uint32_t j = blockIdx.x * 256u + threadIdx.x;
    if (j >= count) return;
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
__global__ void rpatch_kernel(T* __restrict__ out, const uint32_t* __restrict__ idx, const T* __restrict__ err,
                              uint32_t count, uint32_t n, uint32_t* __restrict__ status) {
    const uint32_t j = blockIdx.x * 256u + threadIdx.x;
    if (j >= count) return;
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
    if (__ballot(bad) && lane_id() == 0) atomicOr(status, uint32_t(kNonFiniteOutput));
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
    template void check_finite<T>(const T*, uint32_t, uint32_t*, hipStream_t);
WZ_INST(float)
WZ_INST(double)
#undef WZ_INST

}  // namespace wz
