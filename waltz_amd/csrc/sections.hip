// Side channels (native repairs, quantization outliers) and archive assembly.
//
// Producers append (index, payload) to an unsorted list with wave-aggregated
// atomics and set the index's bit in a row-major flag bitmap. The canonical
// archive needs both lists in increasing index order (cpu/src/codec.cpp:225-257);
// since indices are unique, an entry's sorted position is simply the number of
// flagged indices below it. finish_archive computes that rank from per-1024-bit
// segment popcounts (one exclusive scan) plus in-segment word popcounts and
// scatters every entry straight into its archive section: O(n) work, no sort,
// and the same bytes as a comparison sort.
#include "common.hpp"
#include "launch.hpp"

namespace wz {

namespace {
WZ_D uint32_t lane_id() { return threadIdx.x & 63u; }

// Reserve `want` slots per lane with one atomic per wave.
WZ_D uint32_t wave_reserve(bool want, uint32_t* counter) {
    const unsigned long long mask = __ballot(want);
    if (!mask) return 0;
    const uint32_t leader = uint32_t(__ffsll((long long)mask) - 1);
    uint32_t base = 0;
    if (lane_id() == leader) base = atomicAdd(counter, uint32_t(__popcll(mask)));
    base = __shfl(base, int(leader), 64);
    return base + uint32_t(__popcll(mask & ((1ull << lane_id()) - 1ull)));
}

// Exact |value - original| <= bound with a TwoDiff tie check
// (cpu/src/repair_avx2.cpp:48-56).
WZ_D bool within_bound(double value, double original, double bound) {
    const double high = value - original, magnitude = __builtin_fabs(high);
    if (!__builtin_isfinite(high) || magnitude > bound) return false;
    if (magnitude < bound) return true;
    const double bv = value - high, av = high + bv;
    const double br = bv - original, ar = value - av, low = ar + br;
    return !((high > 0 && low > 0) || (high < 0 && low < 0));
}
}  // namespace

// Native repairs (cpu/src/repair_avx2.cpp:58-74). Each wave covers 64
// consecutive indices starting at a multiple of 64, so it owns two bitmap
// words and writes them without atomics.
template <class T>
__global__ void __launch_bounds__(256) repair_kernel(const T* __restrict__ original, const T* __restrict__ rec,
                                                     uint32_t n, double bound, SideList r,
                                                     uint32_t* __restrict__ status) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    const bool active = i < n;
    bool patch = false, ok = true;
    T err = T(0);
    if (active) {
        T value = rec[i];
        const T o = original[i];
        err = value - o;
        if (!__builtin_isfinite(value) || !__builtin_isfinite(err)) ok = false;
        if (__builtin_fabs(err) > T(bound)) {
            patch = true;
            value = value - err;
        }
        if (!__builtin_isfinite(value) || !within_bound(double(value), double(o), bound)) ok = false;
    }
    const unsigned long long mask = __ballot(patch);
    if (lane_id() == 0 && i < n) r.bitmap[i >> 5] = uint32_t(mask);
    if (lane_id() == 32 && i < n) r.bitmap[i >> 5] = uint32_t(mask >> 32);
    const uint32_t slot = wave_reserve(patch, r.count);
    if (patch && slot < r.capacity) {
        r.index[slot] = i;
        static_cast<T*>(r.payload)[slot] = err;
    }
    if (__ballot(!ok) && lane_id() == 0) atomicOr(status, uint32_t(kRepairBound));
}

template <class T>
void detect_repairs(const T* original, const T* reconstructed, uint32_t n, double bound, SideList r,
                    uint32_t* status, hipStream_t s) {
    hipLaunchKernelGGL((repair_kernel<T>), dim3((n + 255) / 256), dim3(256), 0, s, original, reconstructed, n, bound,
                       r, status);
}

// ------------------------------------------------------------------ ordering
// seg[k] = popcount of bitmap bits [1024k, 1024k+1024) for list `blockIdx.y`.
__global__ void __launch_bounds__(256) segment_count(const uint32_t* __restrict__ bm0,
                                                     const uint32_t* __restrict__ bm1, uint32_t words,
                                                     uint32_t segs, uint32_t* __restrict__ seg0,
                                                     uint32_t* __restrict__ seg1, const uint32_t* cnt0,
                                                     const uint32_t* cnt1) {
    const uint32_t* bm = blockIdx.y ? bm1 : bm0;
    uint32_t* seg = blockIdx.y ? seg1 : seg0;
    if (*(blockIdx.y ? cnt1 : cnt0) == 0) return;
    const uint32_t k = blockIdx.x * 8 + (threadIdx.x >> 5);  // 32 lanes per segment
    if (k >= segs) return;
    const uint32_t w = k * 32 + (threadIdx.x & 31u);
    uint32_t c = w < words ? uint32_t(__builtin_popcount(bm[w])) : 0u;
#pragma unroll
    for (int d = 16; d >= 1; d >>= 1) c += __shfl_xor(c, d, 32);
    if ((threadIdx.x & 31u) == 0) seg[k] = c;
}

// In-place exclusive scan of seg[0..segs) for list blockIdx.x (one workgroup each).
__global__ void __launch_bounds__(1024) segment_scan(uint32_t* seg0, uint32_t* seg1, uint32_t segs, const uint32_t* cnt0,
                                                     const uint32_t* cnt1) {
    uint32_t* seg = blockIdx.x ? seg1 : seg0;
    if (*(blockIdx.x ? cnt1 : cnt0) == 0) return;
    __shared__ uint32_t wsum[16];
    __shared__ uint32_t carry_sh;
    if (threadIdx.x == 0) carry_sh = 0;
    __syncthreads();
    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
    for (uint32_t base = 0; base < segs; base += 1024) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t v = i < segs ? seg[i] : 0u;
        uint32_t inc = v;
#pragma unroll
        for (uint32_t d = 1; d < 64; d <<= 1) {
            const uint32_t t = __shfl_up(inc, d, 64);
            if (lane >= d) inc += t;
        }
        if (lane == 63) wsum[wv] = inc;
        __syncthreads();
        uint32_t before = carry_sh, all = 0;
        for (uint32_t k = 0; k < 16; ++k) {
            before += k < wv ? wsum[k] : 0u;
            all += wsum[k];
        }
        if (i < segs) seg[i] = before + inc - v;
        __syncthreads();
        if (threadIdx.x == 0) carry_sh += all;
        __syncthreads();
    }
}

template <class T>
__global__ void layout_kernel(uint8_t* __restrict__ archive, uint64_t capacity, uint32_t n, SideList rp, SideList qo,
                              uint8_t flags, uint8_t lx, uint8_t lz, uint8_t bz, double quantum,
                              EncodeResult* __restrict__ res, uint32_t* __restrict__ status) {
    const uint32_t np = *rp.count, nq = *qo.count;
    uint32_t st = *status;
    if (np > rp.capacity || nq > qo.capacity || (n > (1u << 20) && nq >= (1u << 20)) ||
        (n > (1u << 22) && np >= (1u << 22)))
        st |= kSideCapacity;
    const uint64_t wb = res->wzp_bytes;
    const Layout L = archive_layout(wb, np, uint64_t(np) * sizeof(T), nq);
    if (L.end > capacity) st |= kWzpCapacity;
    res->np = np;
    res->nq = nq;
    res->end = L.end;
    res->status = st;
    *status = st;
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

WZ_D uint32_t sorted_position(const uint32_t* __restrict__ bm, const uint32_t* __restrict__ seg, uint32_t idx) {
    const uint32_t k = idx >> 10, w = idx >> 5;
    uint32_t pos = seg[k];
    for (uint32_t v = k * 32; v < w; ++v) pos += uint32_t(__builtin_popcount(bm[v]));
    return pos + uint32_t(__builtin_popcount(bm[w] & ((1u << (idx & 31u)) - 1u)));
}

template <class T>
__global__ void __launch_bounds__(256) scatter_kernel(uint8_t* __restrict__ archive, SideList rp, SideList qo,
                                                      const uint32_t* __restrict__ seg0,
                                                      const uint32_t* __restrict__ seg1,
                                                      const EncodeResult* __restrict__ res) {
    if (res->status) return;
    const uint32_t np = res->np, nq = res->nq;
    const Layout L = archive_layout(res->wzp_bytes, np, uint64_t(np) * sizeof(T), nq);
    const uint32_t stride = gridDim.x * 256u;
    for (uint32_t j = blockIdx.x * 256u + threadIdx.x; j < np; j += stride) {
        const uint32_t idx = rp.index[j];
        const uint32_t pos = sorted_position(rp.bitmap, seg0, idx);
        reinterpret_cast<uint32_t*>(archive + L.pi)[pos] = idx;
        reinterpret_cast<T*>(archive + L.pv)[pos] = static_cast<const T*>(rp.payload)[j];
    }
    for (uint32_t j = blockIdx.x * 256u + threadIdx.x; j < nq; j += stride) {
        const uint32_t idx = qo.index[j];
        const uint32_t pos = sorted_position(qo.bitmap, seg1, idx);
        reinterpret_cast<uint32_t*>(archive + L.qi)[pos] = idx;
        reinterpret_cast<uint32_t*>(archive + L.qv)[pos] = static_cast<const uint32_t*>(qo.payload)[j];
    }
}

template <class T>
void finish_archive(uint8_t* archive, uint64_t capacity, uint32_t n, SideList repairs, SideList qout, uint32_t* seg,
                    uint8_t flags, uint8_t lx, uint8_t lz, uint8_t bz, double quantum, EncodeResult* result,
                    hipStream_t s) {
    const uint32_t words = (n + 31) / 32, segs = (n + 1023) / 1024;
    uint32_t* seg0 = seg;
    uint32_t* seg1 = seg + segs;
    hipLaunchKernelGGL(segment_count, dim3((segs + 7) / 8, 2), dim3(256), 0, s, repairs.bitmap, qout.bitmap, words,
                       segs, seg0, seg1, repairs.count, qout.count);
    hipLaunchKernelGGL(segment_scan, dim3(2), dim3(1024), 0, s, seg0, seg1, segs, repairs.count, qout.count);
    hipLaunchKernelGGL((layout_kernel<T>), dim3(1), dim3(1), 0, s, archive, capacity, n, repairs, qout, flags, lx, lz,
                       bz, quantum, result, &result->status);
    hipLaunchKernelGGL((scatter_kernel<T>), dim3(1024), dim3(256), 0, s, archive, repairs, qout, seg0, seg1, result);
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
    out[i] = out[i] - e;  // cpu/src/codec.cpp:330
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

template <class T>
void check_finite(const T* out, uint32_t n, uint32_t* status, hipStream_t s) {
    hipLaunchKernelGGL((finite_kernel<T>), dim3((n + 255) / 256), dim3(256), 0, s, out, n, status);
}

#define WZ_INST(T)                                                                                                \
    template void detect_repairs<T>(const T*, const T*, uint32_t, double, SideList, uint32_t*, hipStream_t);        \
    template void finish_archive<T>(uint8_t*, uint64_t, uint32_t, SideList, SideList, uint32_t*, uint8_t, uint8_t, \
                                    uint8_t, uint8_t, double, EncodeResult*, hipStream_t);                         \
    template void apply_qoutliers<T>(T*, const uint8_t*, const uint8_t*, uint32_t, uint32_t, double, uint32_t*,    \
                                     hipStream_t);                                                                \
    template void apply_repairs<T>(T*, const uint8_t*, const uint8_t*, uint32_t, uint32_t, uint32_t*, hipStream_t); \
    template void check_finite<T>(const T*, uint32_t, uint32_t*, hipStream_t);
WZ_INST(float)
WZ_INST(double)
#undef WZ_INST

}  // namespace wz
