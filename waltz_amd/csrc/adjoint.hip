// Adjoints (transposes) of the multi-level CDF 9/7 transforms, for gradients:
// DWT^T maps a coefficient-space gradient to sample space, IDWT^T maps a
// sample-space gradient to coefficient space. No reference counterpart.
//
// One axis pass of the forward transform lifts every line of the active box:
// the line is extended by whole-sample symmetric reflection, the four lifting
// steps and the scaling run on the extended sequence, and the result is
// restricted to [0, n) and split (even samples first, odd after). The lifting
// steps keep the reflection symmetry, so the pass equals
// split . restrict . L . reflect with L the lifting of an unbounded line. Its
// transpose is reflect^T . L^T . restrict^T . split^T: interleave the split
// gradient, extend it by zeros, run the transposed steps in reverse order
// (each "x[k] += c (x[k-1] + x[k+1])" on one parity becomes the same update on
// the other parity), and fold every extended position back onto the sample it
// reflects to. The inverse pass is the mirror image (split input, natural
// output). The support of L^T is 4 samples wide on each side, so a line chunk
// of 64 outputs is computed from a zero-extended window with 16 samples of
// margin staged in LDS, one 8-sample tile per thread in registers (from the
// tile's 16-sample neighbourhood), with no barrier between the updates.
//
// The adjoint of a multi-level transform runs the transposed passes in the
// reverse order of the transform's schedule.
#include "common.hpp"
#include "launch.hpp"

namespace wz {
namespace {

constexpr int kChunkOut = 64;                  // outputs per line chunk
constexpr int kMargin = 16;                    // window margin on each side
constexpr int kWin = kChunkOut + 2 * kMargin;  // 96 window samples

WZ_HD uint32_t split_index(uint32_t p, uint32_t n) { return (p >> 1) + ((p & 1u) ? half_len(n) : 0u); }

// Lines per workgroup: X lines are rows (window along x); Y and Z lines are
// 32 consecutive x, so every global access is a 128-byte (f32) row segment.
constexpr int kLines = 32;

// Transposed lifting of one 8-sample tile (window positions [i0, i0 + 8))
// from its 16-sample neighbourhood [i0 - 4, i0 + 12) in registers: the four
// updates, each "x[k] += c (x[k-1] + x[k+1])" on one parity, lose one pair of
// validity at the tile ends, so pairs 2..5 of the 8 loaded are final. Every
// value is the same sequence of operations as the same update run over the
// whole window.
// (`ws`, `os`: element strides of the neighbourhood and the result).
template <class T, bool OfInverse>
WZ_D void tile_transposed(const T* w, T* out, int ws = 1, int os = 1) {
    using C = Lift<T>;
    T e[8], o[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        e[j] = w[(2 * j) * ws];
        o[j] = w[(2 * j + 1) * ws];
    }
    if constexpr (OfInverse) {  // transposes of A', B', G', D' in that order
#pragma unroll
        for (int j = 1; j < 8; ++j) e[j] = fmadd(-C::a, o[j - 1] + o[j], e[j]);
#pragma unroll
        for (int j = 1; j < 7; ++j) o[j] = fmadd(-C::b, e[j] + e[j + 1], o[j]);
#pragma unroll
        for (int j = 2; j < 7; ++j) e[j] = fmadd(-C::g, o[j - 1] + o[j], e[j]);
#pragma unroll
        for (int j = 2; j < 6; ++j) o[j] = fmadd(-C::d, e[j] + e[j + 1], o[j]);
    } else {  // transposes of D, G, B, A in that order (after the scaling)
#pragma unroll
        for (int j = 0; j < 7; ++j) o[j] = fmadd(C::d, e[j] + e[j + 1], o[j]);
#pragma unroll
        for (int j = 1; j < 7; ++j) e[j] = fmadd(C::g, o[j - 1] + o[j], e[j]);
#pragma unroll
        for (int j = 1; j < 6; ++j) o[j] = fmadd(C::b, e[j] + e[j + 1], o[j]);
#pragma unroll
        for (int j = 2; j < 6; ++j) e[j] = fmadd(C::a, o[j - 1] + o[j], e[j]);
    }
#pragma unroll
    for (int j = 2; j < 6; ++j) {
        out[(2 * (j - 2)) * os] = e[j];
        out[(2 * (j - 2) + 1) * os] = o[j];
    }
}

// OfInverse: transpose of the inverse pass (natural -> split), else of the
// forward pass (split -> natural).
template <class T, bool OfInverse, int Axis>
__global__ void __launch_bounds__(256) adjoint_pass(const T* __restrict__ src, T* __restrict__ dst, Dims full,
                                                    Dims act, uint32_t chunks, uint32_t groups) {
    constexpr int L = kLines;
    __shared__ T v[L][kWin + 1];
    // transposed-lifting results at positions [c0 - 8, c0 + 72): the chunk
    // and the extended positions the fold reads
    constexpr int kRes = kChunkOut + 16;
    __shared__ T r[L][kRes + 1];
    __shared__ uint64_t lbase[L];
    using C = Lift<T>;
    const uint32_t n = Axis == 0 ? act.nx : Axis == 1 ? act.ny : act.nz;
    const uint32_t chunk = blockIdx.x % chunks, rest = blockIdx.x / chunks;
    const uint32_t group = rest % groups, outer = rest / groups;
    const uint32_t c0 = chunk * kChunkOut;
    const uint64_t plane = uint64_t(full.nx) * full.ny;
    const uint64_t stride = Axis == 0 ? 1u : Axis == 1 ? uint64_t(full.nx) : plane;

    // line l of this workgroup: base offset, or ~0 when out of the box
    auto line_base = [&](uint32_t l) -> uint64_t {
        if constexpr (Axis == 0) {
            const uint64_t li = uint64_t(group) * L + l;
            if (li >= uint64_t(act.ny) * act.nz) return ~uint64_t(0);
            const uint32_t y = uint32_t(li % act.ny), z = uint32_t(li / act.ny);
            return uint64_t(z) * plane + uint64_t(y) * full.nx;
        } else {
            const uint32_t x = group * L + l;
            if (x >= act.nx) return ~uint64_t(0);
            return Axis == 1 ? uint64_t(outer) * plane + x : uint64_t(outer) * full.nx + x;
        }
    };

    if (threadIdx.x < uint32_t(L)) lbase[threadIdx.x] = line_base(threadIdx.x);
    __syncthreads();

    // zero-extended window [c0 - margin, c0 + 64 + margin) of the interleaved line
    for (uint32_t w = threadIdx.x; w < uint32_t(L * kWin); w += 256) {
        const uint32_t l = Axis == 0 ? w / kWin : w % L, i = Axis == 0 ? w % kWin : w / L;
        const int32_t q = int32_t(c0) - kMargin + int32_t(i);
        const uint64_t base = lbase[l];
        T x = T(0);
        if (base != ~uint64_t(0) && q >= 0 && uint32_t(q) < n) {
            const uint32_t pos = OfInverse ? uint32_t(q) : split_index(uint32_t(q), n);
            x = src[base + uint64_t(pos) * stride];
        }
        if (!OfInverse) x = (i & 1u) ? x * (-C::ik) : x * C::k;  // scaling first
        v[l][i] = x;
    }
    __syncthreads();
    // the chunk's 8 tiles of every line, one per thread; the tile before it
    // (images -p, p = 1..4, at the line start) and the one after it (images
    // 2n - 2 - p at the line end) only where those exist
    static_assert(L * (kChunkOut / 8) == 256, "one chunk tile per thread");
    auto tile = [&](uint32_t l, int k) {  // window positions [kMargin + 8k, +8)
        tile_transposed<T, OfInverse>(&v[l][kMargin + 8 * k - 4], &r[l][8 + 8 * k]);
    };
    tile(threadIdx.x % L, int(threadIdx.x / L));
    if (c0 == 0 && threadIdx.x < uint32_t(L)) tile(threadIdx.x, -1);
    if (c0 + kChunkOut + 8 > n && threadIdx.x >= uint32_t(L) && threadIdx.x < uint32_t(2 * L))
        tile(threadIdx.x - L, kChunkOut / 8);
    __syncthreads();

    // fold the extended positions onto the samples they reflect to
    const int32_t lo = -4, hi = int32_t(n) + 4;  // support of the transposed lifting
    for (uint32_t w = threadIdx.x; w < uint32_t(L * kChunkOut); w += 256) {
        const uint32_t l = Axis == 0 ? w / kChunkOut : w % L, i = Axis == 0 ? w % kChunkOut : w / L;
        const uint32_t p = c0 + i;
        const uint64_t base = lbase[l];
        if (base == ~uint64_t(0) || p >= n) continue;
        auto u = [&](int32_t q) { return r[l][q - int32_t(c0) + 8]; };
        T s = u(int32_t(p));
        if (n >= 9) {  // at most two more images: -p and 2n-2-p
            if (p >= 1 && p <= 4) s += u(-int32_t(p));
            const int32_t m = 2 * int32_t(n) - 2 - int32_t(p);
            if (m != int32_t(p) && m < hi) s += u(m);
        } else {  // short line: the whole window is one chunk
            for (int32_t q = lo; q < hi; ++q)
                if (q != int32_t(p) && reflect_pos(q, n) == p) s += u(q);
        }
        if (OfInverse) s = (p & 1u) ? s * (-C::k) : s * C::ik;
        dst[base + uint64_t(OfInverse ? split_index(p, n) : p) * stride] = s;
    }
}

// The transposes of the inverse's X and Y passes (Y^T, then X^T) of one z
// plane in one launch: a workgroup owns 64 x 32 outputs (natural positions
// [x0, x0 + 64) x [y0, y0 + 32)) and stages the zero-extended input window
// [x0 - 12, x0 + 76) x [y0 - 12, y0 + 44) in LDS. Y^T runs as 8-sample
// register tiles on all 88 columns of rows [y0 - 8, y0 + 40), its fold gives
// rows [y0, y0 + 32) (columns outside the box stay zero, as X^T's
// zero extension needs), X^T then runs on those rows over [x0 - 8, x0 + 72),
// and its fold writes the outputs at their split positions in both axes.
// Each value is the same sequence of operations as in the two separate
// passes. Lines of at least 9 samples (the fold of adjoint_pass's long-line
// case).
constexpr int kXYOx = 64, kXYOy = 32;                 // outputs per tile
constexpr int kXYWx = kXYOx + 24, kXYWy = kXYOy + 24;  // input window
template <class T>
__global__ void __launch_bounds__(256) adjoint_xy_kernel(const T* __restrict__ src, T* __restrict__ dst, Dims full,
                                                         Dims act, uint32_t tx, uint32_t ty) {
    __shared__ T A[kXYWy][kXYWx + 1];      // window; then the Y-folded rows (rows 0..31)
    __shared__ T B[kXYOy + 16][kXYWx + 1];  // Y^T results of rows [y0 - 8, y0 + 40); then X^T results
    using C = Lift<T>;
    const uint32_t bx = blockIdx.x % tx, rest = blockIdx.x / tx, by = rest % ty, z = rest / ty;
    const int32_t x0 = int32_t(bx) * kXYOx, y0 = int32_t(by) * kXYOy;
    const int32_t nx = int32_t(act.nx), ny = int32_t(act.ny);
    const uint64_t plane = uint64_t(full.nx) * full.ny;
    const T* sp = src + uint64_t(z) * plane;
    const uint32_t t = threadIdx.x;
    for (uint32_t w = t; w < uint32_t(kXYWy * kXYWx); w += 256) {
        const int32_t i = int32_t(w / kXYWx), j = int32_t(w % kXYWx), y = y0 - 12 + i, x = x0 - 12 + j;
        A[i][j] = y >= 0 && y < ny && x >= 0 && x < nx ? sp[uint64_t(y) * full.nx + uint32_t(x)] : T(0);
    }
    __syncthreads();
    // Y^T: column j, tile k -> rows [y0 - 8 + 8k, +8) from rows [y0 - 12 + 8k, +16)
    for (uint32_t w = t; w < uint32_t(6 * kXYWx); w += 256) {
        const uint32_t j = w % kXYWx, k = w / kXYWx;
        tile_transposed<T, true>(&A[8 * k][j], &B[8 * k][j], kXYWx + 1, kXYWx + 1);
    }
    __syncthreads();
    for (uint32_t w = t; w < uint32_t(kXYOy * kXYWx); w += 256) {
        const int32_t i = int32_t(w / kXYWx), j = int32_t(w % kXYWx), p = y0 + i;
        T v = T(0);
        if (p < ny) {
            auto u = [&](int32_t q) { return B[q - y0 + 8][j]; };
            v = u(p);
            if (p >= 1 && p <= 4) v += u(-p);
            const int32_t m = 2 * ny - 2 - p;
            if (m != p && m < ny + 4) v += u(m);
            v = (p & 1) ? v * (-C::k) : v * C::ik;
        }
        A[i][j] = v;
    }
    __syncthreads();
    // X^T: row i, tile k -> columns [x0 - 8 + 8k, +8) from [x0 - 12 + 8k, +16)
    for (uint32_t w = t; w < uint32_t(10 * kXYOy); w += 256) {
        const uint32_t i = w % kXYOy, k = w / kXYOy;
        tile_transposed<T, true>(&A[i][8 * k], &B[i][8 * k]);
    }
    __syncthreads();
    T* dp = dst + uint64_t(z) * plane;
    for (uint32_t w = t; w < uint32_t(kXYOy * kXYOx); w += 256) {
        const int32_t i = int32_t(w / kXYOx), x = x0 + int32_t(w % kXYOx), p = y0 + i;
        if (x >= nx || p >= ny) continue;
        auto u = [&](int32_t q) { return B[i][q - x0 + 8]; };
        T v = u(x);
        if (x >= 1 && x <= 4) v += u(-x);
        const int32_t m = 2 * nx - 2 - x;
        if (m != x && m < nx + 4) v += u(m);
        v = (x & 1) ? v * (-C::k) : v * C::ik;
        dp[uint64_t(split_index(uint32_t(p), act.ny)) * full.nx + split_index(uint32_t(x), act.nx)] = v;
    }
}

// Only for planes up to 128 x 128: there the launch saved beats the halo
// (each workgroup stages 2.4x its outputs); on a 256 x 256 plane the two
// passes took 82 µs and the joint launch 100 µs.
bool adjoint_xy_fits(Dims act) {
    return act.nx >= 9 && act.ny >= 9 && act.nz && uint64_t(act.nx) * act.ny <= 128u * 128u;
}

template <class T>
void launch_adjoint_xy(const T* src, T* dst, Dims full, Dims act, hipStream_t s) {
    const uint32_t tx = (act.nx + kXYOx - 1) / kXYOx, ty = (act.ny + kXYOy - 1) / kXYOy;
    const uint64_t blocks = uint64_t(tx) * ty * act.nz;
    check_launch_size(blocks * 256);
    hipLaunchKernelGGL((adjoint_xy_kernel<T>), dim3(uint32_t(blocks)), dim3(256), 0, s, src, dst, full, act, tx, ty);
}

Dims shrink_box(Dims d, bool xy, bool z) {
    if (xy) {
        d.nx = half_len(d.nx);
        d.ny = half_len(d.ny);
    }
    if (z) d.nz = half_len(d.nz);
    return d;
}

template <class T, bool OfInverse, int Axis>
void launch_adjoint(const T* src, T* dst, Dims full, Dims act, hipStream_t s) {
    constexpr int L = kLines;
    const uint32_t n = Axis == 0 ? act.nx : Axis == 1 ? act.ny : act.nz;
    if (!n || !act.nx || !act.ny || !act.nz) return;
    const uint32_t chunks = (n + kChunkOut - 1) / kChunkOut;
    const uint64_t groups =
        Axis == 0 ? (uint64_t(act.ny) * act.nz + L - 1) / L : (uint64_t(act.nx) + L - 1) / L;
    const uint64_t outer = Axis == 0 ? 1u : Axis == 1 ? act.nz : act.ny;
    const uint64_t blocks = uint64_t(chunks) * groups * outer;
    check_launch_size(blocks * 256);
    hipLaunchKernelGGL((adjoint_pass<T, OfInverse, Axis>), dim3(uint32_t(blocks)), dim3(256), 0, s, src, dst, full,
                       act, chunks, uint32_t(groups));
}

template <class T, bool OfInverse>
void adjoint_axis(const T* src, T* dst, Dims full, Dims act, int axis, hipStream_t s) {
    if (axis == 0) launch_adjoint<T, OfInverse, 0>(src, dst, full, act, s);
    else if (axis == 1) launch_adjoint<T, OfInverse, 1>(src, dst, full, act, s);
    else launch_adjoint<T, OfInverse, 2>(src, dst, full, act, s);
}

template <class T>
__global__ void __launch_bounds__(256) adjoint_copy(const T* __restrict__ src, T* __restrict__ dst, Dims full,
                                                    Dims act, uint32_t total) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= total) return;
    const uint32_t x = i % act.nx, r = i / act.nx, y = r % act.ny, z = r / act.ny;
    const uint64_t at = (uint64_t(z) * full.ny + y) * full.nx + x;
    dst[at] = src[at];
}

// One group of passes over one box, out of place through tmp0/tmp1, ending in
// `home`; the first pass reads `first` when given (the whole volume's data).
template <class T, bool OfInverse>
void box_passes(T* home, T* tmp0, T* tmp1, Dims full, Dims act, const int* axes, int count, hipStream_t s,
                const T* first = nullptr) {
    const T* src = first ? first : home;
    for (int k = 0; k < count; ++k) {
        // the transposes of the inverse's X and Y passes, Y^T then X^T: one launch
        if (OfInverse && k + 1 < count && axes[k] == 1 && axes[k + 1] == 0 && adjoint_xy_fits(act)) {
            T* dst = k + 1 == count - 1 ? home : ((k + 1) & 1 ? tmp1 : tmp0);
            if (dst != src) {
                launch_adjoint_xy<T>(src, dst, full, act, s);
                src = dst;
                ++k;
                continue;
            }
        }
        T* dst = k == count - 1 ? home : (k & 1 ? tmp1 : tmp0);
        if (dst == src) dst = tmp0;
        adjoint_axis<T, OfInverse>(src, dst, full, act, axes[k], s);
        src = dst;
    }
    if (src != home) {
        const uint64_t total = uint64_t(act.nx) * act.ny * act.nz;
        check_launch_size(total);
        hipLaunchKernelGGL((adjoint_copy<T>), dim3(uint32_t((total + 255) / 256)), dim3(256), 0, s, src, home, full,
                           act, uint32_t(total));
    }
}
}  // namespace

// In place on `data` (full volume), or from `src` into `data` when src is
// given; tmp0/tmp1: two full-volume scratch arrays. Schedules are those of
// transform_forward / transform_inverse, reversed.
template <class T>
void transform_adjoint(T* data, T* tmp0, T* tmp1, Dims full, unsigned lxy, unsigned lz, bool dyadic,
                       bool of_inverse, hipStream_t s, const T* src) {
    // the adjoint of the inverse starts with a pass over the whole volume,
    // which reads src itself; the other schedules start inside a box
    const T* first = src != data && of_inverse && (lxy || lz) ? src : nullptr;
    if (src && src != data && !first)
        hip_check(hipMemcpyAsync(data, src, uint64_t(full.nx) * full.ny * full.nz * sizeof(T),
                                 hipMemcpyDeviceToDevice, s),
                  "copy");
    if (dyadic) {
        Dims sizes[5], act = full;
        for (unsigned l = 0; l < lxy; ++l) {
            sizes[l] = act;
            act = shrink_box(act, true, true);
        }
        if (!of_inverse) {  // forward: levels 0..L-1, axes Y, X, Z -> reversed
            static const int zxy[3] = {2, 0, 1};
            for (int l = int(lxy) - 1; l >= 0; --l) box_passes<T, false>(data, tmp0, tmp1, full, sizes[l], zxy, 3, s);
        } else {  // inverse: levels L-1..0, axes Z, X, Y -> reversed
            static const int yxz[3] = {1, 0, 2};
            for (unsigned l = 0; l < lxy; ++l)
                box_passes<T, true>(data, tmp0, tmp1, full, sizes[l], yxz, 3, s, l == 0 ? first : nullptr);
        }
        return;
    }
    Dims xy[5], zs[5], act = full;
    for (unsigned l = 0; l < lxy; ++l) {
        xy[l] = act;
        act = shrink_box(act, true, false);
    }
    act = full;
    for (unsigned l = 0; l < lz; ++l) {
        zs[l] = act;
        act = shrink_box(act, false, true);
    }
    static const int zonly[1] = {2};
    if (!of_inverse) {  // forward: Z levels 0..lz-1, then XY levels (Y, X) 0..lxy-1 -> reversed
        static const int xy_rev[2] = {0, 1};
        for (int l = int(lxy) - 1; l >= 0; --l) box_passes<T, false>(data, tmp0, tmp1, full, xy[l], xy_rev, 2, s);
        for (int l = int(lz) - 1; l >= 0; --l) box_passes<T, false>(data, tmp0, tmp1, full, zs[l], zonly, 1, s);
    } else {  // inverse: XY levels (X, Y) lxy-1..0, then Z levels lz-1..0 -> reversed
        static const int yx[2] = {1, 0};
        for (unsigned l = 0; l < lz; ++l)
            box_passes<T, true>(data, tmp0, tmp1, full, zs[l], zonly, 1, s, l == 0 ? first : nullptr);
        for (unsigned l = 0; l < lxy; ++l)
            box_passes<T, true>(data, tmp0, tmp1, full, xy[l], yx, 2, s, l == 0 && !lz ? first : nullptr);
    }
}

template void transform_adjoint<float>(float*, float*, float*, Dims, unsigned, unsigned, bool, bool, hipStream_t,
                                       const float*);
template void transform_adjoint<double>(double*, double*, double*, Dims, unsigned, unsigned, bool, bool,
                                        hipStream_t, const double*);

}  // namespace wz
