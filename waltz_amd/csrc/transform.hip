// CDF 9/7 lifting passes for gfx950.
//
// One launch lifts one axis of the active low box [0,act.nx)x[0,act.ny)x[0,act.nz)
// of a full row-major volume (x fastest), out of place: src -> dst, both with
// the full volume's strides. Each thread owns one canonical tile of one line:
// it loads P = Tile/2+4 (even, odd) sample pairs into registers (reflected at
// the axis ends), runs the lifting steps there and writes Tile outputs.
//
// Tiles follow the reference CPU halo schedule (cpu/src/transform.cpp:113-179,
// 377-456). Forward tiles and the inverse X tile are bit-neutral; the inverse
// Y tile (8) and Z tile (16 for f32, 8 for f64) are part of the format because
// the two endpoint sums of a Y/Z tile are contracted into one fma each
// (cpu/src/transform.cpp:115-129).
//
// Y and Z lines map consecutive lanes to consecutive x, so every register load
// is a fully coalesced 256 B (f32) / 512 B (f64) wave access; X lines map lanes
// to consecutive tiles of one row (overlapping 64 B windows served from L1).
#include <cmath>
#include <cstdlib>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <type_traits>

#include "common.hpp"
#include "fused.hpp"
#include "launch.hpp"
#include "lifting.hpp"
#include "profile.hpp"

namespace wz {

// Axis: 0 = X, 1 = Y, 2 = Z.
template <class T, bool Inverse, int Axis, int Tile>
__global__ void __launch_bounds__(256) axis_pass(const T* __restrict__ src, T* __restrict__ dst, Dims full,
                                                 Dims act, uint32_t ntiles, uint32_t total,
                                                 uint32_t* __restrict__ status) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= total) return;
    uint32_t x = 0, y = 0, z = 0, t;
    if constexpr (Axis == 0) {
        t = i % ntiles;
        const uint32_t r = i / ntiles;
        y = r % act.ny;
        z = r / act.ny;
    } else if constexpr (Axis == 1) {
        x = i % act.nx;
        const uint32_t r = i / act.nx;
        t = r % ntiles;
        z = r / ntiles;
    } else {
        x = i % act.nx;
        const uint32_t r = i / act.nx;
        y = r % act.ny;
        t = r / act.ny;
    }
    const uint64_t plane = uint64_t(full.nx) * full.ny;
    const uint32_t n = Axis == 0 ? act.nx : Axis == 1 ? act.ny : act.nz;
    const uint64_t stride = Axis == 0 ? 1u : Axis == 1 ? uint64_t(full.nx) : plane;
    const uint64_t base = uint64_t(z) * plane + uint64_t(y) * full.nx + x;
    lift_tile<T, Inverse, Axis, Tile>(src, base, stride, dst, base, stride, n, t, status);
}

// Two consecutive in-plane passes (X then Y, or Y then X) of a small box in
// one launch: one workgroup per z plane stages the plane in LDS, runs the
// first pass LDS to LDS and the second LDS to dst. The same tile lifting as
// axis_pass, so the same bits.
template <class T, bool Inverse, int A0, int A1, uint32_t kPlaneMax>
__global__ void __launch_bounds__(256) plane_pair_pass(const T* __restrict__ src, T* __restrict__ dst, Dims full,
                                                       Dims act) {
    __shared__ T P0[kPlaneMax], P1[kPlaneMax];  // kPlaneMax: samples of a staged plane
    const uint32_t z = blockIdx.x, t = threadIdx.x;
    const uint64_t plane = uint64_t(full.nx) * full.ny, zb = uint64_t(z) * plane;
    const uint32_t cells = act.nx * act.ny;
    for (uint32_t i = t; i < cells; i += 256u) P0[i] = src[zb + uint64_t(i / act.nx) * full.nx + i % act.nx];
    __syncthreads();
    auto pass = [&](auto axis_c, const T* s, uint64_t sbase, uint64_t spitch, T* d, uint64_t dbase, uint64_t dpitch) {
        constexpr int A = decltype(axis_c)::value;
        const uint32_t n = A == 0 ? act.nx : act.ny, lines = A == 0 ? act.ny : act.nx;
        const uint32_t ntiles = (n + 7) / 8;
        for (uint32_t i = t; i < lines * ntiles; i += 256u) {
            const uint32_t tile = i % ntiles, line = i / ntiles;
            // X: line y, stride 1; Y: line x, stride = the row pitch
            const uint64_t so = A == 0 ? uint64_t(line) * spitch : line, ss = A == 0 ? 1u : spitch;
            const uint64_t dof = A == 0 ? uint64_t(line) * dpitch : line, ds = A == 0 ? 1u : dpitch;
            lift_tile<T, Inverse, A, 8>(s, sbase + so, ss, d, dbase + dof, ds, n, tile, nullptr);
        }
    };
    pass(std::integral_constant<int, A0>{}, P0, 0, act.nx, P1, 0, act.nx);
    __syncthreads();
    pass(std::integral_constant<int, A1>{}, P1, 0, act.nx, dst, zb, full.nx);
}

// Box `act` of a full-pitch volume, quantized and dequantized (the value a
// coefficient holds after quantization), to a compact array of pitch act.nx.
template <class T>
__global__ void __launch_bounds__(256) box_requant(const T* __restrict__ src, T* __restrict__ dst, Dims full, Dims act,
                                                   uint32_t total, T q, T inv) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= total) return;
    const uint32_t x = i % act.nx, r = i / act.nx, y = r % act.ny, z = r / act.ny;
    dst[i] = requantize<T>(src[(uint64_t(z) * full.ny + y) * full.nx + x], q, inv);
}

template <class T>
__global__ void __launch_bounds__(256) box_copy(const T* __restrict__ src, T* __restrict__ dst, Dims full, Dims act,
                                                uint32_t total) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= total) return;
    const uint32_t x = i % act.nx, r = i / act.nx, y = r % act.ny, z = r / act.ny;
    const uint64_t at = (uint64_t(z) * full.ny + y) * full.nx + x;
    dst[at] = src[at];
}

namespace {
Dims shrink(Dims d, bool xy, bool z) {
    if (xy) { d.nx = half_len(d.nx); d.ny = half_len(d.ny); }
    if (z) d.nz = half_len(d.nz);
    return d;
}

template <class T, bool Inverse, int Axis, int Tile>
void launch_axis(const T* src, T* dst, Dims full, Dims act, uint32_t* status, hipStream_t stream) {
    const uint32_t n = Axis == 0 ? act.nx : Axis == 1 ? act.ny : act.nz;
    const uint32_t ntiles = (n + Tile - 1) / Tile;
    const uint64_t lines = uint64_t(act.nx) * act.ny * act.nz / n;
    const uint64_t total = lines * ntiles;
    if (total == 0) return;
    check_launch_size(total);
    hipLaunchKernelGGL((axis_pass<T, Inverse, Axis, Tile>), dim3(uint32_t((total + 255) / 256)), dim3(256), 0, stream,
                       src, dst, full, act, ntiles, uint32_t(total), status);
}

// The line tile is free for forward passes and the inverse X pass (it only
// decides how a line is cut among threads; every tile recomputes its halo);
// the inverse Y and Z tiles are the canonical ones.
template <class T, bool Inverse, int Axis>
void launch_axis_tile(const T* src, T* dst, Dims full, Dims act, uint32_t tile, uint32_t* status, hipStream_t s) {
    if (tile == 32) launch_axis<T, Inverse, Axis, 32>(src, dst, full, act, status, s);
    else if (tile == 16) launch_axis<T, Inverse, Axis, 16>(src, dst, full, act, status, s);
    else launch_axis<T, Inverse, Axis, 8>(src, dst, full, act, status, s);
}

template <class T, bool Inverse>
void axis(const T* src, T* dst, Dims full, Dims act, int axis, const uint32_t* tiles, uint32_t* status,
          hipStream_t s) {
    constexpr int ZTileInv = sizeof(T) == 4 ? 16 : 8;
    if (axis == 0) launch_axis_tile<T, Inverse, 0>(src, dst, full, act, tiles[0], status, s);
    else if (Inverse && axis == 1) launch_axis<T, Inverse, 1, 8>(src, dst, full, act, status, s);
    else if (axis == 1) launch_axis_tile<T, Inverse, 1>(src, dst, full, act, tiles[1], status, s);
    else if (Inverse) launch_axis<T, Inverse, 2, ZTileInv>(src, dst, full, act, status, s);
    else launch_axis_tile<T, Inverse, 2>(src, dst, full, act, tiles[2], status, s);
}

template <class T, bool Inverse>
bool xy_pair(const T* src, T* dst, Dims full, Dims act, int a0, hipStream_t s);

// X and Y passes of a box in one launch: planes up to `limit` samples (tuned:
// 1024 or 4096, 0 = none) staged whole in LDS by one workgroup, larger ones
// (forward Y then X, inverse X then Y) as 64 x 32 tiles with their halos.
template <class T, bool Inverse>
bool plane_pair(const T* src, T* dst, Dims full, Dims act, int a0, int a1, uint32_t limit, hipStream_t s) {
    const uint64_t cells = uint64_t(act.nx) * act.ny;
    if (act.nz == 0 || a0 + a1 != 1) return false;
    if (cells > limit) return xy_pair<T, Inverse>(src, dst, full, act, a0, s);
    if (cells <= 1024) {
        if (a0 == 0)
            hipLaunchKernelGGL((plane_pair_pass<T, Inverse, 0, 1, 1024>), dim3(act.nz), dim3(256), 0, s, src, dst,
                               full, act);
        else
            hipLaunchKernelGGL((plane_pair_pass<T, Inverse, 1, 0, 1024>), dim3(act.nz), dim3(256), 0, s, src, dst,
                               full, act);
    } else {
        if (a0 == 0)
            hipLaunchKernelGGL((plane_pair_pass<T, Inverse, 0, 1, 4096>), dim3(act.nz), dim3(256), 0, s, src, dst,
                               full, act);
        else
            hipLaunchKernelGGL((plane_pair_pass<T, Inverse, 1, 0, 4096>), dim3(act.nz), dim3(256), 0, s, src, dst,
                               full, act);
    }
    return true;
}

// Every Z level of the plane schedule in one pass (its Z phase): one
// workgroup per row of 64 columns stages the columns, all z, in LDS, runs the
// levels' Z lifting there (lds_pass: the tiles and arithmetic of the Z axis
// passes, so the same bits) and writes them once. src may equal dst; status:
// screen for non-finite samples the forward's input (its first pass) or the
// inverse's output (its last).
constexpr uint32_t kZColBytes = 66560;
constexpr uint32_t kZColumns = 64;
bool zcol_fits(uint32_t nz, unsigned width) { return nz <= kZColBytes / kZColumns / width; }

template <class T, bool Inverse>
__global__ void __launch_bounds__(kLdsThreads) zcol_levels_kernel(const T* src, T* dst, Dims full, uint32_t levels,
                                                                  uint32_t* __restrict__ status) {
    constexpr int ZTile = Inverse ? (sizeof(T) == 4 ? 16 : 8) : 8;
    __shared__ T A[kZColBytes / sizeof(T)];
    const uint32_t t = threadIdx.x;
    const uint32_t xb = (full.nx + kZColumns - 1) / kZColumns;
    const uint32_t y = blockIdx.x / xb, x0 = (blockIdx.x - y * xb) * kZColumns;
    const uint32_t cols = min(kZColumns, full.nx - x0), cells = cols * full.nz;
    const uint64_t plane = uint64_t(full.nx) * full.ny, base = uint64_t(y) * full.nx + x0;
    const FastDiv by_cols(cols);  // cells fit LDS: exact
    // every thread issues kBurst loads before it writes any of them to LDS
    constexpr uint32_t kBurst = 8;
    bool ok = true;
    for (uint32_t i0 = 0; i0 < cells; i0 += kBurst * kLdsThreads) {
        T v[kBurst];
#pragma unroll
        for (uint32_t k = 0; k < kBurst; ++k) {
            const uint32_t i = i0 + k * kLdsThreads + t, r = by_cols.div(i);
            v[k] = i < cells ? src[uint64_t(r) * plane + base + (i - r * cols)] : T(0);
        }
#pragma unroll
        for (uint32_t k = 0; k < kBurst; ++k) {
            const uint32_t i = i0 + k * kLdsThreads + t, r = by_cols.div(i);
            if (i < cells) {
                ok = ok && (Inverse || !status || __builtin_isfinite(v[k]));
                A[r * kZColumns + (i - r * cols)] = v[k];
            }
        }
    }
    if (__ballot(!ok) && (t & 63u) == 0) atomicOr(status, uint32_t(kNonFiniteInput));
    __syncthreads();
    uint32_t ex[5];
    ex[0] = full.nz;
    for (uint32_t l = 1; l < 5; ++l) ex[l] = half_len(ex[l - 1]);
    for (uint32_t k = 0; k < levels; ++k) {
        const uint32_t l = Inverse ? levels - 1 - k : k;
        lds_pass<T, Inverse, 2, ZTile>(A, kZColumns, ex[l], cols);  // ends with a barrier
    }
    bool fin = true;
    for (uint32_t i = t; i < cells; i += kLdsThreads) {
        const uint32_t r = by_cols.div(i), c = i - r * cols;
        const T v = A[r * kZColumns + c];
        if (Inverse && status) fin = fin && __builtin_isfinite(v);
        dst[uint64_t(r) * plane + base + c] = v;
    }
    if (__ballot(!fin) && (t & 63u) == 0) atomicOr(status, uint32_t(kNonFiniteOutput));
}

template <class T, bool Inverse>
void zcol_levels(const T* src, T* dst, Dims full, unsigned levels, uint32_t* status, hipStream_t s) {
    if (!zcol_fits(full.nz, sizeof(T)) || levels > 5) throw std::logic_error("Z phase does not fit LDS");
    const uint64_t blocks = uint64_t(full.ny) * ((full.nx + kZColumns - 1) / kZColumns);
    check_launch_size(blocks * kLdsThreads);
    hipLaunchKernelGGL((zcol_levels_kernel<T, Inverse>), dim3(uint32_t(blocks)), dim3(kLdsThreads), 0, s, src, dst,
                       full, uint32_t(levels), status);
}

// One XY level of the plane schedule per launch: a workgroup of 256 threads
// owns a 64 x 32 tile of one plane, runs the level's first pass (inverse: X,
// forward: Y) over the tile and the halo the second pass reads, keeps the
// results in LDS and runs the second pass from there, so the level reads its
// input once (halos aside) and writes once. Tiles and arithmetic are those of
// lift_tile (inverse Y tile 8 folded, every other tile 8 and unfolded), so
// every value equals the per-axis passes'.
constexpr uint32_t kXYW = 64, kXYH = 32, kXYThreads = 256;

struct XYTile {
    uint32_t x0, y0, z;
};
// Tile of this workgroup: consecutive tiles go to one XCD (workgroups are
// dealt to the eight XCDs in turn), so neighbours' halos meet in its L2.
WZ_D bool xy_tile(Dims n, XYTile& t) {
    const uint32_t tx = (n.nx + kXYW - 1) / kXYW, ty = (n.ny + kXYH - 1) / kXYH;
    const uint32_t total = tx * ty * n.nz, per = (total + 7) / 8;
    const uint32_t id = (blockIdx.x & 7u) * per + (blockIdx.x >> 3);
    if (id >= total) return false;
    t.x0 = (id % tx) * kXYW;
    t.y0 = ((id / tx) % ty) * kXYH;
    t.z = id / (tx * ty);
    return true;
}
uint32_t xy_blocks(Dims n) {
    const uint64_t total = uint64_t((n.nx + kXYW - 1) / kXYW) * ((n.ny + kXYH - 1) / kXYH) * n.nz;
    check_launch_size(8 * ((total + 7) / 8) * kXYThreads);
    return uint32_t(8 * ((total + 7) / 8));
}

template <class T>
struct XYInverse {
    const T* hi;  // the level's coefficients outside its low box
    uint64_t hpy, hpz;
    const T* lo;  // its low box [0, ceil(nx/2)) x [0, ceil(ny/2)): the coarser reconstruction
    uint64_t lpy, lpz;
    T* out;  // the level's reconstruction (natural layout)
    uint64_t opy, opz;
    Dims n;  // level box: nx x ny, every plane
    // q > 0: the coefficients hold unquantized values, quantized and
    // dequantized as they load (the low box too when lo_coeff: the coarsest
    // level, whose low box is the coefficient volume's)
    T q, inv;
    uint32_t lo_coeff;
};

template <class T>
__global__ void __launch_bounds__(kXYThreads) xy_inverse_kernel(XYInverse<T> a) {
    constexpr int TX = 8, P = TX / 2 + 4;  // X tiles
    constexpr uint32_t R = kXYH + 8;       // y positions y0 - 4 .. y0 + 35 (interleaved)
    constexpr uint32_t kTiles = kXYW / TX;
    __shared__ T S[R][kXYW + 1];
    XYTile tl;
    if (!xy_tile(a.n, tl)) return;
    const uint32_t t = threadIdx.x, nx = a.n.nx, ny = a.n.ny, hx = half_len(nx), hy = half_len(ny);
    const T* hiz = a.hi + uint64_t(tl.z) * a.hpz;
    const T* loz = a.lo + uint64_t(tl.z) * a.lpz;
    // X: row r of S is y position y0 - 4 + r, read from its split row
    for (uint32_t item = t; item < R * kTiles; item += kXYThreads) {
        const uint32_t r = item / kTiles, k = item % kTiles, start = tl.x0 + k * TX;
        if (start >= nx) continue;
        const uint32_t sr = split_pos(int32_t(tl.y0) - 4 + int32_t(r), ny);
        const T* hrow = hiz + uint64_t(sr) * a.hpy;
        const T* lrow = loz + uint64_t(sr) * a.lpy;
        const bool lowrow = sr < hy;
        T e[P], o[P];
#pragma unroll
        for (int j = 0; j < P; ++j) {
            const int32_t q = int32_t(start) - 4 + 2 * j;
            const uint32_t ce = split_pos(q, nx), co = split_pos(q + 1, nx);
            const bool le = lowrow && ce < hx, lo_ = lowrow && co < hx;
            e[j] = le ? lrow[ce] : hrow[ce];
            o[j] = lo_ ? lrow[co] : hrow[co];
            if (a.q > T(0)) {  // (wave-uniform)
                if (!le || a.lo_coeff) e[j] = requantize<T>(e[j], a.q, a.inv);
                if (!lo_ || a.lo_coeff) o[j] = requantize<T>(o[j], a.q, a.inv);
            }
        }
        lift_inverse<T, P, false>(e, o);
#pragma unroll
        for (int j = 2; j < P - 2; ++j) {
            const uint32_t q = start + 2u * uint32_t(j - 2);
            if (q < nx) S[r][q - tl.x0] = e[j];
            if (q + 1 < nx) S[r][q + 1 - tl.x0] = o[j];
        }
    }
    __syncthreads();
    // Y (tile 8, folded): one (column, tile) per thread, lanes 64 consecutive x
    constexpr int PY = 8;
    const uint32_t c = t & 63u, k = t >> 6, x = tl.x0 + c, start = tl.y0 + 8 * k;
    if (x >= nx || start >= ny) return;
    T e[PY], o[PY];
#pragma unroll
    for (int j = 0; j < PY; ++j) {
        e[j] = S[8 * k + 2 * j][c];
        o[j] = S[8 * k + 2 * j + 1][c];
    }
    lift_inverse<T, PY, true>(e, o);
    T* col = a.out + uint64_t(tl.z) * a.opz + x;
#pragma unroll
    for (int j = 2; j < PY - 2; ++j) {
        const uint32_t q = start + 2u * uint32_t(j - 2);
        if (q < ny) col[uint64_t(q) * a.opy] = e[j];
        if (q + 1 < ny) col[uint64_t(q + 1) * a.opy] = o[j];
    }
}

template <class T>
struct XYForward {
    const T* src;  // the level's input (natural layout)
    uint64_t spy, spz;
    T* hi;  // where the three high bands go
    uint64_t hpy, hpz;
    T* lo;  // where the low band [0, ceil(nx/2)) x [0, ceil(ny/2)) goes
    uint64_t lpy, lpz;
    Dims n;
    uint32_t* status;  // non-null: screen the input for non-finite samples
};

template <class T>
__global__ void __launch_bounds__(kXYThreads) xy_forward_kernel(XYForward<T> a) {
    constexpr int TL = 8, P = TL / 2 + 4;  // Y and X tiles
    constexpr uint32_t C = kXYW + 8;       // x positions x0 - 4 .. x0 + 67 (reflected)
    __shared__ T S[kXYH][C + 1];
    XYTile tl;
    if (!xy_tile(a.n, tl)) return;
    const uint32_t t = threadIdx.x, nx = a.n.nx, ny = a.n.ny, hx = half_len(nx), hy = half_len(ny);
    const T* srcz = a.src + uint64_t(tl.z) * a.spz;
    // Y: column c of S is x position x0 - 4 + c (its reflected column); row r
    // is the Y output of y position y0 + r (its split row)
    bool ok = true;
    for (uint32_t item = t; item < C * (kXYH / TL); item += kXYThreads) {
        const uint32_t c = item % C, k = item / C, start = tl.y0 + k * TL;
        if (start >= ny) continue;
        const T* col = srcz + reflect_pos(int32_t(tl.x0) - 4 + int32_t(c), nx);
        T e[P], o[P];
#pragma unroll
        for (int j = 0; j < P; ++j) {
            const int32_t q = int32_t(start) - 4 + 2 * j;
            e[j] = col[uint64_t(reflect_pos(q, ny)) * a.spy];
            o[j] = col[uint64_t(reflect_pos(q + 1, ny)) * a.spy];
        }
        if (a.status) {  // input screen (cpu/src/codec.cpp:183)
#pragma unroll
            for (int j = 0; j < P; ++j) ok = ok && finite_value(e[j]) && finite_value(o[j]);
        }
        lift_forward<T, P>(e, o);
#pragma unroll
        for (int j = 2; j < P - 2; ++j) {
            const uint32_t q = start + 2u * uint32_t(j - 2);
            if (q < ny) S[q - tl.y0][c] = e[j];
            if (q + 1 < ny) S[q + 1 - tl.y0][c] = o[j];
        }
    }
    if (a.status && !ok) atomicOr(a.status, uint32_t(kNonFiniteInput));
    __syncthreads();
    // X: one (row, tile) per thread; row r is split row sy of y = y0 + r
    constexpr uint32_t kTiles = kXYW / TL;
    for (uint32_t item = t; item < kXYH * kTiles; item += kXYThreads) {
        const uint32_t r = item / kTiles, k = item % kTiles, y = tl.y0 + r, start = tl.x0 + k * TL;
        if (y >= ny || start >= nx) continue;
        const uint32_t sy = (y & 1u) ? hy + (y >> 1) : (y >> 1);
        T e[P], o[P];
#pragma unroll
        for (int j = 0; j < P; ++j) {
            const uint32_t q = start + 2u * uint32_t(j) - tl.x0;  // S column of x position start - 4 + 2j
            e[j] = S[r][q];
            o[j] = S[r][q + 1];
        }
        lift_forward<T, P>(e, o);
        const bool lowrow = sy < hy;
        T* hrow = a.hi + uint64_t(tl.z) * a.hpz + uint64_t(sy) * a.hpy;
        T* lrow = a.lo + uint64_t(tl.z) * a.lpz + uint64_t(sy) * a.lpy;
#pragma unroll
        for (int j = 2; j < P - 2; ++j) {
            const uint32_t q = start + 2u * uint32_t(j - 2), ce = q >> 1, co = hx + (q >> 1);
            if (q < nx) (lowrow && ce < hx ? lrow : hrow)[ce] = e[j];
            if (q + 1 < nx) hrow[co] = o[j];
        }
    }
}

template <class T>
void xy_inverse(const XYInverse<T>& a, hipStream_t s) {
    hipLaunchKernelGGL((xy_inverse_kernel<T>), dim3(xy_blocks(a.n)), dim3(kXYThreads), 0, s, a);
}
template <class T>
void xy_forward(const XYForward<T>& a, hipStream_t s) {
    hipLaunchKernelGGL((xy_forward_kernel<T>), dim3(xy_blocks(a.n)), dim3(kXYThreads), 0, s, a);
}

// The X and Y passes of box `act` (in full-pitched arrays, src != dst) as one
// tiled launch, in the order the tiled kernels lift (forward Y then X,
// inverse X then Y); false for the other order.
template <class T, bool Inverse>
bool xy_pair(const T* src, T* dst, Dims full, Dims act, int a0, hipStream_t s) {
    const uint64_t py = full.nx, pz = uint64_t(full.nx) * full.ny;
    if constexpr (Inverse) {
        if (a0 != 0) return false;
        xy_inverse<T>(XYInverse<T>{src, py, pz, src, py, pz, dst, py, pz, act, T(0), T(0), 0u}, s);
    } else {
        if (a0 != 1) return false;
        xy_forward<T>(XYForward<T>{src, py, pz, dst, py, pz, dst, py, pz, act, nullptr}, s);
    }
    return true;
}

// Compact XY boxes of the plane schedule's levels >= 1 (every plane) in
// `scratch`: level l's box at offset 0 for odd l and after the largest
// level-1 box for even l (the two never alias).
template <class T>
T* xy_box(T* scratch, Dims full, unsigned l) {
    const uint64_t first = uint64_t(half_len(full.nx)) * half_len(full.ny) * full.nz;
    return (l & 1u) ? scratch : scratch + ((first + 63) & ~uint64_t(63));
}

template <class T>
void copy_box(const T* src, T* dst, Dims full, Dims act, hipStream_t s) {
    const uint64_t total = uint64_t(act.nx) * act.ny * act.nz;
    if (!total) return;
    check_launch_size(total);
    hipLaunchKernelGGL((box_copy<T>), dim3(uint32_t((total + 255) / 256)), dim3(256), 0, s, src, dst, full, act,
                       uint32_t(total));
}

// Runs a sequence of axis passes on one box: data starts in `home` (or in
// `first_src` for the very first pass), every pass is out of place, and the
// box ends in `home` (or in `last_dst` for the very last pass).
template <class T, bool Inverse>
struct Passes {
    T* home;
    T* tmp[2];
    const T* first_src;
    T* last_dst;
    uint32_t* status;
    hipStream_t s;
    Tuning tune;  // tiles and LDS pair limit of the per-axis passes
    bool first = true;
    unsigned base = 0;  // level number of the box's level 0 (stage names)

    // Passes: src -> tmp0 -> tmp1 -> ... -> final. A single pass whose source
    // is already `home` goes through tmp0 and is copied back.
    void level(Dims full, Dims act, const int* axes, int count, bool last_level, unsigned index) {
        static const char* fwd[] = {"forward_L0", "forward_L1", "forward_L2", "forward_L3", "forward_L4"};
        static const char* inv[] = {"inverse_L0", "inverse_L1", "inverse_L2", "inverse_L3", "inverse_L4"};
        StageClock clock(Inverse ? inv[index + base] : fwd[index + base], s);
        const T* src = first ? first_src : home;
        T* final_dst = last_level && last_dst ? last_dst : home;
        for (int k = 0; k < count; ++k) {
            const bool screen = first && k == 0 && status;
            // two in-plane passes of a small box: one launch
            const bool pair = !screen && k + 1 < count && axes[k] != 2 && axes[k + 1] != 2;
            const int last = pair ? k + 1 : k;
            T* const other = src == tmp[0] ? tmp[1] : tmp[0];  // never the source
            T* dst = last == count - 1 ? final_dst : other;
            if (dst == src) dst = other;  // in place at the end: through scratch, copied back below
            if (pair && plane_pair<T, Inverse>(src, dst, full, act, axes[k], axes[k + 1], tune.pair_max, s)) {
                ++k;
            } else {
                axis<T, Inverse>(src, dst, full, act, axes[k], tune.tile, screen ? status : nullptr, s);
            }
            src = dst;
        }
        if (src != final_dst) copy_box<T>(src, final_dst, full, act, s);
        first = false;
    }
};
}  // namespace

// ------------------------------------------------------------------ fused dyadic schedule
namespace {
// z segment per workgroup: enough segments for ~1024 workgroups, a multiple of
// 16 (the canonical f32 inverse Z tile, and twice the f64 one). Bit-neutral.
uint32_t z_segment(Dims n) {
    const uint64_t tiles = uint64_t((n.nx + 63) / 64) * ((n.ny + 31) / 32);
    uint64_t segs = (1024 + tiles - 1) / tiles;
    if (segs < 1) segs = 1;
    uint64_t sz = (n.nz + segs - 1) / segs;
    sz = (sz + 15) & ~uint64_t(15);
    return uint32_t(sz < 16 ? 16 : sz);
}

// Leading levels whose box fills the GPU with fused workgroups (>= 512 of
// them at the smallest z segment); coarser boxes are cheaper as axis passes.
unsigned fused_levels(Dims n, unsigned levels) {
    unsigned f = 0;
    for (; f < levels; ++f) {
        const uint64_t wgs = uint64_t((n.nx + 63) / 64) * ((n.ny + 31) / 32) * ((n.nz + 15) / 16);
        if (wgs < 512) break;
        n = shrink(n, true, true);
    }
    return f;
}

// ------------------------------------------------------------------ tuning
// Launch choices of the dyadic schedule that do not change a single bit: how
// many leading levels run as fused kernels and each fused level's z segment.
// Per (geometry, type, direction): measured by autotune(), set explicitly, or
// the heuristic above.
struct TuneKey {
    uint32_t nx, ny, nz, width, inverse;
    bool operator<(const TuneKey& o) const {
        return std::tie(nx, ny, nz, width, inverse) < std::tie(o.nx, o.ny, o.nz, o.width, o.inverse);
    }
};
std::mutex g_tune_mu;
std::map<TuneKey, Tuning> g_tune;

Tuning heuristic_impl(Dims full, unsigned lxy) {
    Tuning t{};
    t.fused = fused_levels(full, lxy);
    t.tile[0] = t.tile[1] = t.tile[2] = 8;
    t.pair_max = 1024;
    Dims n = full;
    for (unsigned l = 0; l < 5; ++l) {
        t.sz[l] = z_segment(n);
        n = shrink(n, true, true);
    }
    return t;
}
// Compact boxes of levels >= 1 live in `scratch`: level l's box at offset 0
// for odd l and after the largest level-1 box for even l (two never-aliasing
// halves, each <= n/8 + lines).
template <class T>
T* low_box(T* scratch, Dims full, unsigned l) {
    const Dims h{half_len(full.nx), half_len(full.ny), half_len(full.nz)};
    const uint64_t first = uint64_t(h.nx) * h.ny * h.nz;
    return (l & 1u) ? scratch : scratch + ((first + 63) & ~uint64_t(63));
}
}  // namespace

Tuning tuning_for(Dims full, unsigned lxy, unsigned width, bool inverse) {
    std::lock_guard<std::mutex> g(g_tune_mu);
    const auto it = g_tune.find(TuneKey{full.nx, full.ny, full.nz, width, inverse ? 1u : 0u});
    Tuning t = it != g_tune.end() ? it->second : heuristic_impl(full, lxy);
    if (t.fused > lxy) t.fused = lxy;
    if (!fused_plane_fits(full.nx, full.ny, width)) t.fused = 0;  // planes of 4 GiB and more: axis passes
    return t;
}

Tuning heuristic_tuning(Dims full, unsigned lxy) { return heuristic_impl(full, lxy); }

void clear_tuning() {
    std::lock_guard<std::mutex> g(g_tune_mu);
    g_tune.clear();
}

void set_tuning(Dims full, unsigned width, bool inverse, const Tuning& t) {
    for (unsigned l = 0; l < 5; ++l)
        if (t.sz[l] < 16 || t.sz[l] % 16) throw std::invalid_argument("z segment must be a positive multiple of 16");
    if (t.fused > 5) throw std::invalid_argument("at most five fused levels");
    for (unsigned a = 0; a < 3; ++a)
        if (t.tile[a] != 8 && t.tile[a] != 16 && t.tile[a] != 32)
            throw std::invalid_argument("pass tile must be 8, 16 or 32");
    if (t.pair_max != 0 && t.pair_max != 1024 && t.pair_max != 4096)
        throw std::invalid_argument("LDS pair limit must be 0, 1024 or 4096");
    std::lock_guard<std::mutex> g(g_tune_mu);
    g_tune[TuneKey{full.nx, full.ny, full.nz, width, inverse ? 1u : 0u}] = t;
}

std::string tuning_report() {
    std::lock_guard<std::mutex> g(g_tune_mu);
    std::string out = "[";
    for (const auto& [k, t] : g_tune) {
        if (out.size() > 1) out += ",";
        out += "{\"shape_xyz\":[" + std::to_string(k.nx) + "," + std::to_string(k.ny) + "," + std::to_string(k.nz) +
               "],\"width\":" + std::to_string(k.width) + ",\"inverse\":" + (k.inverse ? "true" : "false") +
               ",\"fused_levels\":" + std::to_string(t.fused) + ",\"z_segment\":[";
        for (unsigned l = 0; l < 5; ++l) out += (l ? "," : "") + std::to_string(t.sz[l]);
        out += "],\"pass_tiles_xyz\":[" + std::to_string(t.tile[0]) + "," + std::to_string(t.tile[1]) + "," +
               std::to_string(t.tile[2]) + "],\"lds_pair_max\":" + std::to_string(t.pair_max) + "}";
    }
    return out + "]";
}

template <class T>
void fused_forward(const T* input, T* coeff, T* tmp0, Dims full, unsigned levels, const Tuning& tu, uint32_t* status,
                   hipStream_t s) {
    const T* src = input;
    uint64_t spy = full.nx, spz = uint64_t(full.nx) * full.ny;
    Dims n = full;
    static const char* names[] = {"forward_L0", "forward_L1", "forward_L2", "forward_L3", "forward_L4"};
    for (unsigned l = 0; l < levels; ++l) {
        StageClock clock(names[l], s);
        const Dims h{half_len(n.nx), half_len(n.ny), half_len(n.nz)};
        FwdLevel<T> a;
        a.src = src;
        a.spy = spy;
        a.spz = spz;
        a.n = n;
        a.hi = coeff;
        a.hpy = full.nx;
        a.hpz = uint64_t(full.nx) * full.ny;
        if (l + 1 == levels) {
            a.lo = coeff;
            a.lpy = a.hpy;
            a.lpz = a.hpz;
        } else {
            a.lo = low_box(tmp0, full, l + 1);
            a.lpy = h.nx;
            a.lpz = uint64_t(h.nx) * h.ny;
        }
        a.sz = tu.sz[l];
        a.status = l == 0 ? status : nullptr;
        launch_fwd_level<T>(a, s);
        src = a.lo;
        spy = a.lpy;
        spz = a.lpz;
        n = h;
    }
}

// The finest inverse level as a fused kernel: high bands from `coeff`, the
// coarser reconstruction from `lo` (pitches lpy, lpz), into `output`, with
// the caller's emitter (store, screen and folded repairs, or the repair
// emitter, which may requantize the high bands on load).
template <class T>
void launch_finest(const T* coeff, const T* lo, uint32_t lpy, uint32_t lpz, T* output, Dims full, uint32_t sz,
                   const Level0Emit<T>* emit, hipStream_t s, const char* name) {
    const bool repairs = emit && emit->mode == kEmitRepair;
    const bool patch = emit && emit->mode == kEmitStoreScreen && emit->np;
    const uint64_t keys = patch ? patch_table_words(full, emit->np, sizeof(T)) : 0;
    if (!keys && patch) throw std::logic_error("repair list cannot be folded into the finest level");
    if (patch && !emit->ptable_built && emit->plan) throw std::logic_error("planned repair keys not built");
    if (patch && !emit->ptable_built) {  // the key table of the repairs the finest level applies
        StageClock clock("patch_keys", s);
        repair_keys<T>(emit->pidx, emit->perr, emit->np, full, emit->ptable, emit->status, s);
    }
    StageClock clock(repairs ? "inverse_L0_repairs" : name, s);
    InvLevel<T> a{};
    a.hi = coeff;
    a.hpy = full.nx;
    a.hpz = uint64_t(full.nx) * full.ny;
    a.lo = lo;
    a.lpy = lpy;
    a.lpz = lpz;
    a.n = full;
    a.out = output;
    a.opy = full.nx;
    a.opz = uint64_t(full.nx) * full.ny;
    a.sz = sz;
    a.emit = kEmitStore;
    if (emit) {
        a.emit = emit->mode;
        a.status = emit->status;
        a.orig = emit->orig;
        a.bound = emit->bound;
        if (repairs && emit->quantum > 0) {
            a.q = T(emit->quantum);
            a.inv = T(1.0 / emit->quantum);
        }
        a.bound_lo = float(emit->bound);
        if (double(a.bound_lo) > emit->bound) a.bound_lo = std::nextafter(a.bound_lo, 0.0f);
        a.rep = emit->rep;
        if (patch) {
            a.pidx = emit->pidx;
            a.perr = emit->perr;
            a.ptable = emit->ptable;
            a.np = emit->np;
            a.keys = uint32_t(keys);
            a.plan = emit->plan;
            a.archive = emit->archive;
        }
        if (repairs) {  // the segment layout the caller allocated fixes the z segment
            a.sz = emit->rep.cap / (64u * 32u);
            if (a.sz % 16 || inv_segments(full, a.sz) != emit->rep.segments)
                throw std::logic_error("repair segments do not match the finest level's launch");
        }
        if (repairs && (full.nx & 31u))  // atomics path needs a clean bitmap
            zero_words(emit->rep.bitmap, (uint64_t(full.nx) * full.ny * full.nz + 31) / 32, s);
    }
    launch_inv_level<T>(a, s);
}

// Fused levels levels-1 .. 1 (the coarse LLL read from coeff): where their
// reconstruction of the level-1 box lies (coeff itself when only the finest
// level is fused).
template <class T>
LowBox<T> fused_coarse(T* coeff, T* tmp0, Dims full, unsigned levels, const Tuning& tu, hipStream_t s,
                       unsigned base) {
    Dims sizes[5], n = full;
    for (unsigned l = 0; l < levels; ++l) { sizes[l] = n; n = shrink(n, true, true); }
    LowBox<T> low{coeff, full.nx, full.nx * full.ny};
    static const char* names[] = {"inverse_L0", "inverse_L1", "inverse_L2", "inverse_L3", "inverse_L4"};
    for (int l = int(levels) - 1; l >= 1; --l) {
        StageClock clock(names[l + base], s);
        InvLevel<T> a{};
        a.hi = coeff;
        a.hpy = full.nx;
        a.hpz = uint64_t(full.nx) * full.ny;
        a.lo = low.p;
        a.lpy = low.py;
        a.lpz = low.pz;
        a.n = sizes[l];
        a.out = low_box(tmp0, full, unsigned(l));
        a.opy = sizes[l].nx;
        a.opz = uint64_t(sizes[l].nx) * sizes[l].ny;
        a.sz = tu.sz[l];
        a.emit = kEmitStore;
        launch_inv_level<T>(a, s);
        low = LowBox<T>{a.out, a.opy, a.opz};
    }
    return low;
}

template <class T>
void fused_inverse(T* coeff, T* output, T* tmp0, Dims full, unsigned levels, const Tuning& tu,
                   const Level0Emit<T>* emit, hipStream_t s, unsigned base) {
    static const char* names[] = {"inverse_L0", "inverse_L1", "inverse_L2", "inverse_L3", "inverse_L4"};
    const LowBox<T> low = fused_coarse<T>(coeff, tmp0, full, levels, tu, s, base);
    launch_finest<T>(coeff, low.p, low.py, low.pz, output, full, tu.sz[0], emit, s, names[base]);
}

void level0_repair_segments(Dims full, unsigned lxy, bool dyadic, unsigned width, uint32_t* segments,
                            uint32_t* cap) {
    *segments = 0;
    *cap = 0;
    if (!dyadic || !lxy) return;
    const Tuning tu = tuning_for(full, lxy, width, true);
    if (!tu.fused) return;
    const uint32_t sz = width == 4 ? repair_z_segment<float>(full, tu.sz[0]) : repair_z_segment<double>(full, tu.sz[0]);
    *segments = uint32_t(inv_segments(full, sz));
    *cap = 64u * 32u * sz;
}

// Full multi-level transform (cpu/src/transform.cpp:409-456).
// Forward: reads `input` (may equal coeff), leaves coefficients in `coeff`.
// Inverse: reads coefficients in `coeff` (overwritten), writes the last pass
// into `output` (may equal coeff). tmp0/tmp1: two full-volume scratch arrays.
template <class T>
void transform_forward(const T* input, T* coeff, T* tmp0, T* tmp1, Dims full, unsigned lxy, unsigned lz,
                       bool dyadic, uint32_t* status, hipStream_t s) {
    Passes<T, false> p{coeff, {tmp0, tmp1}, input, nullptr, status, s, tuning_for(full, lxy, sizeof(T), false)};
    if (lxy == 0 && lz == 0) {
        if (input != coeff) copy_box<T>(input, coeff, full, full, s);
        return;
    }
    if (dyadic) {
        const Tuning tu = tuning_for(full, lxy, sizeof(T), false);
        const unsigned fused = tu.fused;
        Dims act = full;
        if (fused) {
            if (input == coeff) {  // the level-0 kernel must not read what it writes
                copy_box<T>(input, tmp1, full, full, s);
                input = tmp1;
            }
            fused_forward<T>(input, coeff, tmp0, full, fused, tu, status, s);  // leaves LLL in coeff
            p.first = false;
            for (unsigned l = 0; l < fused; ++l) act = shrink(act, true, true);
        }
        static const int yxz[3] = {1, 0, 2};
        for (unsigned l = fused; l < lxy; ++l) {
            p.level(full, act, yxz, 3, false, l);
            act = shrink(act, true, true);
        }
        return;
    }
    static const int zonly[1] = {2};
    static const int yx[2] = {1, 0};
    // every XY level as one fused launch: the Z phase (one LDS-staged pass)
    // into tmp1, then level l reads its box (tmp1, then the compact boxes in
    // tmp0) and writes its high bands into coeff and its low band into the
    // next box (coeff for the last level)
    if (lxy && (lz ? zcol_fits(full.nz, sizeof(T)) : input != coeff)) {
        const T* src = input;
        uint64_t spy = full.nx, spz = uint64_t(full.nx) * full.ny;
        if (lz) {
            StageClock clock("forward_Z", s);
            zcol_levels<T, false>(input, tmp1, full, lz, status, s);
            src = tmp1;
        }
        Dims n = full;
        static const char* names[] = {"forward_L0", "forward_L1", "forward_L2", "forward_L3", "forward_L4"};
        for (unsigned l = 0; l < lxy; ++l) {
            StageClock clock(names[l], s);
            const Dims h{half_len(n.nx), half_len(n.ny), n.nz};
            const bool last = l + 1 == lxy;
            XYForward<T> a{src, spy, spz, coeff, full.nx, uint64_t(full.nx) * full.ny, nullptr, 0, 0, n,
                           l == 0 && !lz ? status : nullptr};
            a.lo = last ? coeff : xy_box(tmp0, full, l + 1);
            a.lpy = last ? full.nx : h.nx;
            a.lpz = last ? uint64_t(full.nx) * full.ny : uint64_t(h.nx) * h.ny;
            xy_forward<T>(a, s);
            src = a.lo;
            spy = a.lpy;
            spz = a.lpz;
            n = h;
        }
        return;
    }
    Dims act = full;
    if (lz && zcol_fits(full.nz, sizeof(T))) {  // the Z phase in one LDS-staged pass
        StageClock clock("forward_Z", s);
        zcol_levels<T, false>(p.first_src, coeff, full, lz, status, s);
        p.first = false;
    } else {
        for (unsigned l = 0; l < lz; ++l) {
            p.level(full, act, zonly, 1, false, l);
            act = shrink(act, false, true);
        }
    }
    act = full;
    for (unsigned l = 0; l < lxy; ++l) {
        p.level(full, act, yx, 2, false, l);
        act = shrink(act, true, false);
    }
}

namespace {
// The inverse with launch choices `tu`; `base`: the level number of this
// field's level 0 (a coarse box run as a field of its own), for stage names.
template <class T>
bool inverse_impl(T* coeff, T* output, T* tmp0, T* tmp1, Dims full, unsigned lxy, unsigned lz, bool dyadic,
                  hipStream_t s, const Level0Emit<T>* emit, const Tuning& tu, unsigned base, double requant = 0.0) {
    if (requant > 0 && (dyadic || !plane_fused_fits(full, lxy, lz, sizeof(T))))
        throw std::logic_error("requantizing inverse without fused XY levels");
    Passes<T, true> p{coeff, {tmp0, tmp1}, coeff, output == coeff ? nullptr : output, nullptr, s, tu};
    p.base = base;
    if (lxy == 0 && lz == 0) {
        if (output != coeff) copy_box<T>(coeff, output, full, full, s);
        return false;
    }
    if (dyadic) {
        const unsigned fused = tu.fused;
        Dims sizes[5], act = full;
        for (unsigned l = 0; l < lxy; ++l) { sizes[l] = act; act = shrink(act, true, true); }
        static const int zxy[3] = {2, 0, 1};
        // coarse levels as passes in coeff; then the fused levels (coarse LLL read from coeff)
        for (int l = int(lxy) - 1; l >= int(fused); --l)
            p.level(full, sizes[l], zxy, 3, !fused && l == 0, unsigned(l));
        if (!fused) return false;
        if (output == coeff) {  // the level-0 kernel must not overwrite what it reads
            fused_inverse<T>(coeff, tmp1, tmp0, full, fused, tu, emit, s, base);
            if (!emit || emit->mode != kEmitRepair) copy_box<T>(tmp1, coeff, full, full, s);
        } else {
            fused_inverse<T>(coeff, output, tmp0, full, fused, tu, emit, s, base);
        }
        return emit != nullptr;
    }
    Dims xy[5], zs[5], act = full;
    for (unsigned l = 0; l < lxy; ++l) { xy[l] = act; act = shrink(act, true, false); }
    act = full;
    for (unsigned l = 0; l < lz; ++l) { zs[l] = act; act = shrink(act, false, true); }
    static const int xyo[2] = {0, 1};
    static const int zonly[1] = {2};
    // every XY level as one fused launch: level l reads coeff's bands and the
    // coarser reconstruction (coeff's low quadrant for the coarsest level,
    // then the compact boxes in tmp0) and writes its box to tmp0, the finest
    // into tmp1, from which the Z phase (one LDS-staged pass) writes output
    if (lxy && (lz ? zcol_fits(full.nz, sizeof(T)) : output != coeff)) {
        // the Z phase screens the output when asked (no repairs to fold here)
        uint32_t* screen = emit && emit->mode == kEmitStoreScreen && !emit->np && lz ? emit->status : nullptr;
        static const char* names[] = {"inverse_L0", "inverse_L1", "inverse_L2", "inverse_L3", "inverse_L4"};
        const T* lo = coeff;
        uint64_t lpy = full.nx, lpz = uint64_t(full.nx) * full.ny;
        for (int l = int(lxy) - 1; l >= 0; --l) {
            StageClock clock(names[l + base], s);
            XYInverse<T> a{coeff, full.nx, uint64_t(full.nx) * full.ny, lo, lpy, lpz, nullptr, 0, 0, xy[l],
                           T(requant), requant > 0 ? T(1.0 / requant) : T(0), uint32_t(lo == coeff)};
            if (l == 0) {
                a.out = lz ? tmp1 : output;
                a.opy = full.nx;
                a.opz = uint64_t(full.nx) * full.ny;
            } else {
                a.out = xy_box(tmp0, full, unsigned(l));
                a.opy = xy[l].nx;
                a.opz = uint64_t(xy[l].nx) * xy[l].ny;
            }
            xy_inverse<T>(a, s);
            lo = a.out;
            lpy = a.opy;
            lpz = a.opz;
        }
        if (lz) {
            StageClock clock("inverse_Z", s);
            zcol_levels<T, true>(tmp1, output, full, lz, screen, s);
        }
        return screen != nullptr;
    }
    for (int l = int(lxy) - 1; l >= 0; --l) p.level(full, xy[l], xyo, 2, lz == 0 && l == 0, unsigned(l));
    if (lz && zcol_fits(full.nz, sizeof(T))) {  // the Z phase in one LDS-staged pass
        StageClock clock("inverse_Z", s);
        zcol_levels<T, true>(p.first ? p.first_src : p.home, output, full, lz, nullptr, s);
    } else {
        for (int l = int(lz) - 1; l >= 0; --l) p.level(full, zs[l], zonly, 1, l == 0, unsigned(l));
    }
    return false;
}

uint64_t pad64(uint64_t v) { return (v + 63) & ~uint64_t(63); }
uint64_t volume(Dims d) { return uint64_t(d.nx) * d.ny * d.nz; }
}  // namespace

template <class T>
bool transform_inverse(T* coeff, T* output, T* tmp0, T* tmp1, Dims full, unsigned lxy, unsigned lz, bool dyadic,
                       hipStream_t s, const Level0Emit<T>* emit, double requant) {
    return inverse_impl<T>(coeff, output, tmp0, tmp1, full, lxy, lz, dyadic, s, emit,
                           tuning_for(full, lxy, sizeof(T), true), 0, requant);
}

template <class T>
LowBox<T> inverse_coarse_levels(T* coeff, T* tmp0, T* tmp1, Dims full, unsigned lxy, hipStream_t s) {
    const Tuning tu = tuning_for(full, lxy, sizeof(T), true);
    if (!lxy || !tu.fused) throw std::logic_error("inverse split without a fused finest level");
    Passes<T, true> p{coeff, {tmp0, tmp1}, coeff, nullptr, nullptr, s, tu};
    Dims sizes[5], act = full;
    for (unsigned l = 0; l < lxy; ++l) { sizes[l] = act; act = shrink(act, true, true); }
    static const int zxy[3] = {2, 0, 1};
    for (int l = int(lxy) - 1; l >= int(tu.fused); --l) p.level(full, sizes[l], zxy, 3, false, unsigned(l));
    return fused_coarse<T>(coeff, tmp0, full, tu.fused, tu, s, 0);
}

template <class T>
void inverse_finest_level(const T* coeff, const LowBox<T>& low, T* output, Dims full, unsigned lxy,
                          const Level0Emit<T>& emit, hipStream_t s) {
    if (output == coeff) throw std::logic_error("the finest level must not overwrite its coefficients");
    const Tuning tu = tuning_for(full, lxy, sizeof(T), true);
    launch_finest<T>(coeff, low.p, low.py, low.pz, output, full, tu.sz[0], &emit, s, "inverse_L0");
}

template <class T>
bool transform_inverse_preserving(const T* coeff, T* output, T* tmp0, T* tmp1, Dims full, unsigned lxy,
                                  hipStream_t s) {
    const Tuning tu = tuning_for(full, lxy, sizeof(T), true);
    if (!lxy || !tu.fused || output == coeff) return false;
    const Dims h1 = shrink(full, true, true);
    copy_box<T>(coeff, output, full, h1, s);
    LowBox<T> low = inverse_coarse_levels<T>(output, tmp0, tmp1, full, lxy, s);
    if (low.p == output) {  // the finest level's coarse input would be overwritten as it stores
        copy_box<T>(output, tmp1, full, h1, s);
        low.p = tmp1;
    }
    const Level0Emit<T> store{kEmitStore, nullptr, nullptr, 0.0, SegList{}, 0.0};
    inverse_finest_level<T>(coeff, low, output, full, lxy, store, s);
    return true;
}

bool plane_fused_fits(Dims full, unsigned lxy, unsigned lz, unsigned width) {
    return lxy && lz && zcol_fits(full.nz, width);
}

bool coarse_split_fits(Dims full, unsigned lxy, unsigned width) {
    if (lxy < 2 || tuning_for(full, lxy, width, true).fused < 2) return false;
    const Dims h1 = shrink(full, true, true), h2 = shrink(h1, true, true);
    return 4 * pad64(volume(h2)) <= volume(full) && volume(h1) <= volume(full);
}

template <class T>
void inverse_coarse_requant(const T* coeff, T* t0, T* t1, Dims full, unsigned lxy, double quantum, hipStream_t s) {
    if (!coarse_split_fits(full, lxy, sizeof(T))) throw std::logic_error("coarse inverse split does not apply");
    const Tuning tu = tuning_for(full, lxy, sizeof(T), true);
    const Dims h1 = shrink(full, true, true), h2 = shrink(h1, true, true);
    const uint64_t v2 = pad64(volume(h2));
    T *c2 = t1, *o2 = t1 + v2, *w0 = t1 + 2 * v2, *w1 = t1 + 3 * v2;
    const T q = T(quantum), inv = T(1.0 / quantum);
    {
        StageClock clock("coarse_requant", s);
        const uint64_t total = volume(h2);
        check_launch_size(total);
        hipLaunchKernelGGL((box_requant<T>), dim3(uint32_t((total + 255) / 256)), dim3(256), 0, s, coeff, c2, full, h2,
                           uint32_t(total), q, inv);
    }
    const T* lo = c2;
    if (lxy > 2) {  // levels >= 2: the box as a field of its own, with this field's launch choices
        Tuning t2 = tu;
        t2.fused = tu.fused - 2;
        for (unsigned l = 0; l < 5; ++l) t2.sz[l] = l + 2 < 5 ? tu.sz[l + 2] : 16u;
        inverse_impl<T>(c2, o2, w0, w1, h2, lxy - 2, lxy - 2, true, s, nullptr, t2, 2);
        lo = o2;
    }
    StageClock clock("inverse_L1", s);
    InvLevel<T> a{};
    a.hi = coeff;
    a.hpy = full.nx;
    a.hpz = uint64_t(full.nx) * full.ny;
    a.lo = lo;
    a.lpy = h2.nx;
    a.lpz = uint64_t(h2.nx) * h2.ny;
    a.n = h1;
    a.out = t0;
    a.opy = h1.nx;
    a.opz = uint64_t(h1.nx) * h1.ny;
    a.sz = tu.sz[1];
    a.emit = kEmitStore;
    a.q = q;
    a.inv = inv;
    launch_inv_level<T>(a, s);
}

template <class T>
void inverse_finest_requant(const T* coeff, const T* t0, Dims full, unsigned lxy, const Level0Emit<T>& emit,
                            hipStream_t s) {
    const Tuning tu = tuning_for(full, lxy, sizeof(T), true);
    const Dims h1 = shrink(full, true, true);
    launch_finest<T>(coeff, t0, h1.nx, h1.nx * h1.ny, nullptr, full, tu.sz[0], &emit, s, "inverse_L0");
}

template bool transform_inverse_preserving<float>(const float*, float*, float*, float*, Dims, unsigned, hipStream_t);
template bool transform_inverse_preserving<double>(const double*, double*, double*, double*, Dims, unsigned,
                                                   hipStream_t);
template LowBox<float> inverse_coarse_levels<float>(float*, float*, float*, Dims, unsigned, hipStream_t);
template LowBox<double> inverse_coarse_levels<double>(double*, double*, double*, Dims, unsigned, hipStream_t);
template void inverse_finest_level<float>(const float*, const LowBox<float>&, float*, Dims, unsigned,
                                          const Level0Emit<float>&, hipStream_t);
template void inverse_finest_level<double>(const double*, const LowBox<double>&, double*, Dims, unsigned,
                                           const Level0Emit<double>&, hipStream_t);
template void transform_forward<float>(const float*, float*, float*, float*, Dims, unsigned, unsigned, bool,
                                       uint32_t*, hipStream_t);
template void transform_forward<double>(const double*, double*, double*, double*, Dims, unsigned, unsigned, bool,
                                        uint32_t*, hipStream_t);
template bool transform_inverse<float>(float*, float*, float*, float*, Dims, unsigned, unsigned, bool, hipStream_t,
                                       const Level0Emit<float>*, double);
template bool transform_inverse<double>(double*, double*, double*, double*, Dims, unsigned, unsigned, bool,
                                        hipStream_t, const Level0Emit<double>*, double);
template void inverse_coarse_requant<float>(const float*, float*, float*, Dims, unsigned, double, hipStream_t);
template void inverse_coarse_requant<double>(const double*, double*, double*, Dims, unsigned, double, hipStream_t);
template void inverse_finest_requant<float>(const float*, const float*, Dims, unsigned, const Level0Emit<float>&,
                                            hipStream_t);
template void inverse_finest_requant<double>(const double*, const double*, Dims, unsigned, const Level0Emit<double>&,
                                             hipStream_t);

}  // namespace wz
