// Side-channel pieces shared by the single-field and batched pipelines:
// repair detection over row-major tiles, index-order ranks from flag bitmaps,
// the archive layout and header, and placement of both side lists (see
// sections.hip for the scheme).
#pragma once

#include "common.hpp"
#include "launch.hpp"

namespace wz {

namespace sec {
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

}  // namespace sec

// Native repairs (cpu/src/repair_avx2.cpp:58-74) over row-major tiles of
// kRepairTile indices (one segment of the SegList each). Round k of tile t
// covers indices t*4096 + 256k + [0,256) (one coalesced load per wave per
// round); each wave owns the two flag-bitmap words of its 64 indices, and the
// tile's entries are compacted in index order through ballot ranks and a
// 64-entry scan in LDS.
template <class T>
WZ_D void repair_tile(const T* __restrict__ original, const T* __restrict__ rec, uint32_t n, double bound,
                      const SegList& r, uint32_t tile, uint32_t* __restrict__ status) {
    constexpr uint32_t rounds = kRepairTile / 256;
    __shared__ uint32_t cnt[rounds * 4];
    const uint64_t tile0 = uint64_t(tile) * kRepairTile;
    const uint32_t lane = sec::lane_id(), wv = threadIdx.x >> 6;
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
        const uint32_t inc = sec::wave_inclusive(v);
        cnt[lane] = inc - v;
        if (lane == 63) r.count[tile] = inc;
    }
    __syncthreads();
    uint32_t* idx = r.index + uint64_t(tile) * r.cap;
    T* val = static_cast<T*>(r.payload) + uint64_t(tile) * r.cap;
#pragma unroll
    for (uint32_t k = 0; k < rounds; ++k)
        if ((ballots[k] >> lane) & 1ull) {
            const uint32_t at = cnt[4 * k + wv] + uint32_t(__popcll(ballots[k] & below));
            idx[at] = uint32_t(tile0 + 256u * k + threadIdx.x);
            val[at] = err[k];
        }
    if (__ballot(!ok) && lane == 0) atomicOr(status, uint32_t(kRepairBound));
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
    uint32_t run = sec::block_exclusive256(sum, sh, total);
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
        const uint32_t inc = sec::wave_inclusive(v);
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
WZ_D void scan_layout(uint8_t* __restrict__ archive, uint64_t capacity, uint32_t n, uint32_t* rbsum, uint32_t* qbsum,
                      uint32_t blocks, uint32_t* np_dev, const SideList& qo, uint8_t flags, uint8_t lx, uint8_t lz,
                      uint8_t bz, double quantum, EncodeResult* __restrict__ res) {
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
                        uint32_t k, uint32_t slice, uint32_t slices = kPlaceSlices) {
    const uint32_t c = rp.count[k];
    if (!c) return;
    const uint32_t np = res->np;
    const Layout L = archive_layout(res->wzp_bytes, np, uint64_t(np) * sizeof(T), res->nq);
    const uint32_t* idx = rp.index + uint64_t(k) * rp.cap;
    const T* val = static_cast<const T*>(rp.payload) + uint64_t(k) * rp.cap;
    uint32_t* oi = reinterpret_cast<uint32_t*>(archive + L.pi);
    T* ov = reinterpret_cast<T*>(archive + L.pv);
    for (uint32_t j = slice * 256u + threadIdx.x; j < c; j += 256u * slices) {
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


}  // namespace wz
